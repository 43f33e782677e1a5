// Covariance and correlation of an n x m data matrix (n observations of m
// variables): the column statistics and the centring / normalising pass,
// then the m x m product dataᵀ·data on the triangle contraction kernels.
//
// Replaces the kernel sequences of polybench/covariance.cpp:76-128 (mean,
// centre, upper-triangle covariance + mirror) and
// polybench/correlation.cpp:92-158 (mean, standard deviation, centre and
// scale, upper-triangle correlation + mirror, unit diagonal).
//
// Column sums run as a two-pass reduction: a grid of (256-column strip x row
// slab) blocks writes per-slab partial sums to the workspace, a second kernel
// adds the slabs of each column in slab order, so results do not depend on
// scheduling.  Every sum is float, as in the reference, made compensated
// (Neumaier) so its error does not grow with the row count.  The arithmetic
// of each finished statistic follows the reference's types: the covariance
// mean divides a float sum by a float constant; the correlation standard
// deviation divides the float sum of float squares by the double constant
// FLOAT_N, takes a float square root and replaces values <= eps by 1;
// centring subtracts in float, and the correlation scaling divides by
// sqrt(FLOAT_N) in double and by the deviation in float.  Those two scalar
// divisions are the only double operations, and they are the reference's own.
#include <algorithm>
#include <cmath>

#include "mfx_internal.h"

namespace mfx {

mfx_status sym_common(Handle* h, mfx_precision prec, mfx_uplo uplo, int mirror, int64_t n, int64_t k, float alpha,
    const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int nprod, bool trans,
    hipStream_t s);
mfx_status check_matrix(const char* name, const void* p, int64_t rows, int64_t cols, int64_t ld);

namespace {

// Compensated (Neumaier) float summation: `sum` plus the running rounding
// error `comp`, so a long column sum stays as accurate as a float can hold
// without widening the type the reference sums in.
struct FSum {
  float sum = 0.0f, comp = 0.0f;
  __device__ __forceinline__ void add(float x) {
    const float t = sum + x;
    comp += fabsf(sum) >= fabsf(x) ? (sum - t) + x : (x - t) + sum;
    sum = t;
  }
  __device__ __forceinline__ float value() const { return sum + comp; }
};

// Partial column sums of rows [slab * rows_per, ...): mode 0 sums x, mode 1
// sums (x - mean[j])^2 with the difference and square in float.
__global__ __launch_bounds__(256) void column_partial_kernel(int mode, int64_t m, int64_t n, const float* data,
    int64_t ld, const float* mean, int64_t rows_per, float* partial) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(j >= m) return;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per, r1 = min(n, r0 + rows_per);
  const float mu = mode == 1 ? mean[j] : 0.0f;
  FSum acc;
  int64_t r = r0;
  // four independent loads in flight per thread
  for(; r + 4 <= r1; r += 4) {
    float x[4];
#pragma unroll
    for(int u = 0; u < 4; ++u) x[u] = data[(r + u) * ld + j];
#pragma unroll
    for(int u = 0; u < 4; ++u) {
      if(mode == 1) {
        const float d = x[u] - mu;
        acc.add(d * d);
      } else {
        acc.add(x[u]);
      }
    }
  }
  for(; r < r1; ++r) {
    const float x = data[r * ld + j];
    if(mode == 1) {
      const float d = x - mu;
      acc.add(d * d);
    } else {
      acc.add(x);
    }
  }
  partial[(int64_t)blockIdx.y * m + j] = acc.value();
}

// mean[j] = (sum over slabs) / n_const, in float.
__global__ void finish_mean_kernel(int64_t m, int slabs, const float* partial, float n_const, float* mean) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= m) return;
  FSum s;
  for(int q = 0; q < slabs; ++q) s.add(partial[(int64_t)q * m + j]);
  mean[j] = s.value() / n_const;
}

// stddev[j] = sqrtf(sum / n_const) with the division in double (the
// reference's FLOAT_N is a double constant), <= eps -> 1.
__global__ void finish_stddev_kernel(int64_t m, int slabs, const float* partial, double n_const, double eps,
    float* stddev) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= m) return;
  FSum s;
  for(int q = 0; q < slabs; ++q) s.add(partial[(int64_t)q * m + j]);
  float v = (float)((double)s.value() / n_const);
  v = sqrtf(v);
  stddev[j] = (double)v <= eps ? 1.0f : v;
}

// out[i][j] = data[i][j] - mean[j]; with a deviation, then divided in double
// by scale_d and in float by stddev[j].  Columns across threads, rows across
// the grid's y dimension.
__global__ __launch_bounds__(256) void center_kernel(int64_t m, int64_t n, const float* data, int64_t ld,
    const float* mean, const float* stddev, double scale_d, float* out, int64_t ldo) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(j >= m) return;
  const float mu = mean[j], sd = stddev ? stddev[j] : 1.0f;
  for(int64_t i = blockIdx.y; i < n; i += gridDim.y) {
    float v = data[i * ld + j] - mu;
    if(stddev) {
      v = (float)((double)v / scale_d);
      v = v / sd;
    }
    out[i * ldo + j] = v;
  }
}

__global__ void set_diagonal_kernel(int64_t m, float* C, int64_t ldc, float value) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(j < m) C[j * ldc + j] = value;
}

// Slabs so that the partial-sum grid has about 2048 blocks.
int64_t slab_count(int64_t m, int64_t n) {
  const int64_t strips = ceil_div(m, 256);
  return std::max<int64_t>(1, std::min<int64_t>(n, ceil_div(2048, strips)));
}

mfx_status column_stats(Handle* h, int mode, int64_t m, int64_t n, const float* data, int64_t ld, const float* mean,
    int64_t* slabs_out, float** partial_out, hipStream_t s) {
  const int64_t slabs = slab_count(m, n);
  void* ws = nullptr;
  MFX_CHECK(workspace(h, (size_t)(slabs * m) * sizeof(float), &ws));
  float* partial = static_cast<float*>(ws);
  const int64_t rows_per = ceil_div(n, slabs);
  MFX_CHECK(timed_launch(h, mode ? "column_sqdev_partial" : "column_sum_partial", s, [&] {
    column_partial_kernel<<<dim3((unsigned)ceil_div(m, 256), (unsigned)slabs), 256, 0, s>>>(mode, m, n, data, ld,
        mean, rows_per, partial);
  }));
  *slabs_out = slabs;
  *partial_out = partial;
  return MFX_SUCCESS;
}

mfx_status check_stats_args(mfx_precision prec, int64_t m, int64_t n, const float* data, int64_t ld,
    const float* centered, int64_t ldw, const float* symmat, int64_t lds) {
  if(prec < MFX_PREC_FP32 || prec > MFX_PREC_BF16) return unsupported("covariance / correlation: FP32, SPLIT_BF16 or BF16");
  if(m <= 0 || n <= 0) return invalid("covariance / correlation need m > 0 and n > 0");
  MFX_CHECK(check_matrix("data", data, n, m, ld));
  MFX_CHECK(check_matrix("centered", centered, n, m, ldw));
  MFX_CHECK(check_matrix("symmat", symmat, m, m, lds));
  return MFX_SUCCESS;
}

}  // namespace
}  // namespace mfx

using namespace mfx;

extern "C" {

mfx_status mfx_covariance(mfx_handle handle, mfx_precision prec, int64_t m, int64_t n, const float* data, int64_t ld,
    float n_const, float* mean, float* centered, int64_t ldw, float* symmat, int64_t lds, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  MFX_CHECK(check_stats_args(prec, m, n, data, ld, centered, ldw, symmat, lds));
  if(!mean) return invalid("null mean");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t slabs = 0;
  float* partial = nullptr;
  MFX_CHECK(column_stats(h, 0, m, n, data, ld, nullptr, &slabs, &partial, s));
  MFX_CHECK(timed_launch(h, "column_mean", s, [&] {
    finish_mean_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, s>>>(m, (int)slabs, partial, n_const, mean);
  }));
  MFX_CHECK(timed_launch(h, "center_columns", s, [&] {
    center_kernel<<<dim3((unsigned)ceil_div(m, 256), (unsigned)std::min<int64_t>(n, 4096)), 256, 0, s>>>(m, n, data,
        ld, mean, nullptr, 1.0, centered, ldw);
  }));
  // symmat = centeredᵀ · centered: upper triangle, mirrored (covariance.cpp:114-126)
  return sym_common(h, prec, MFX_UPPER, 1, m, n, 1.0f, centered, ldw, nullptr, 0, 0.0f, symmat, lds, 1, true, s);
}

mfx_status mfx_correlation(mfx_handle handle, mfx_precision prec, int64_t m, int64_t n, const float* data,
    int64_t ld, double n_const, double eps, float* mean, float* stddev, float* centered, int64_t ldw, float* symmat,
    int64_t lds, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  MFX_CHECK(check_stats_args(prec, m, n, data, ld, centered, ldw, symmat, lds));
  if(!mean || !stddev) return invalid("null mean or stddev");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t slabs = 0;
  float* partial = nullptr;
  MFX_CHECK(column_stats(h, 0, m, n, data, ld, nullptr, &slabs, &partial, s));
  MFX_CHECK(timed_launch(h, "column_mean", s, [&] {
    finish_mean_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, s>>>(m, (int)slabs, partial, (float)n_const, mean);
  }));
  MFX_CHECK(column_stats(h, 1, m, n, data, ld, mean, &slabs, &partial, s));
  MFX_CHECK(timed_launch(h, "column_stddev", s, [&] {
    finish_stddev_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, s>>>(m, (int)slabs, partial, n_const, eps, stddev);
  }));
  MFX_CHECK(timed_launch(h, "center_scale_columns", s, [&] {
    center_kernel<<<dim3((unsigned)ceil_div(m, 256), (unsigned)std::min<int64_t>(n, 4096)), 256, 0, s>>>(m, n, data,
        ld, mean, stddev, std::sqrt(n_const), centered, ldw);
  }));
  // symmat = centeredᵀ · centered: upper triangle, mirrored, then a unit
  // diagonal (correlation.cpp:138-158)
  MFX_CHECK(sym_common(h, prec, MFX_UPPER, 1, m, n, 1.0f, centered, ldw, nullptr, 0, 0.0f, symmat, lds, 1, true, s));
  return timed_launch(h, "unit_diagonal", s, [&] {
    set_diagonal_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, s>>>(m, symmat, lds, 1.0f);
  });
}

}  // extern "C"
