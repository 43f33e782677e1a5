// FP64 oracle contraction and error metrics, plus input generation.
//
// The FP64 GEMM is a register-tiled VALU kernel (64x64 tile, 4x4 outputs per
// thread, sequential k with fma): FP64 is used on the GPU only here, to score
// the FP32 / split-BF16 / BF16 / FP8 products, never inside a product.
// Reductions run over a fixed grid with a fixed tree so results are
// deterministic run to run.
#include "mfx_internal.h"

namespace mfx {
namespace {

constexpr int T64 = 64, K64 = 16;

struct F64Params {
  int M, N, K;
  double alpha, beta;
  const double* A; int64_t lda; int ta;
  const double* B; int64_t ldb; int tb;
  double* C; int64_t ldc;
};

__global__ __launch_bounds__(256) void gemm_f64_kernel(const F64Params p) {
  __shared__ double As[K64][T64 + 1];
  __shared__ double Bs[K64][T64 + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int m0 = blockIdx.y * T64, n0 = blockIdx.x * T64;
  double acc[4][4];
#pragma unroll
  for(int i = 0; i < 4; ++i)
#pragma unroll
    for(int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for(int k0 = 0; k0 < p.K; k0 += K64) {
#pragma unroll
    for(int e = 0; e < 4; ++e) {
      const int idx = tid + e * 256;  // 1024 = 64 x 16
      // A panel: (m, k) with m fastest when A is transposed, k fastest otherwise
      int m, k;
      if(p.ta) { k = idx / T64; m = idx % T64; } else { m = idx / K64; k = idx % K64; }
      const int gm = m0 + m, gk = k0 + k;
      double va = 0.0;
      if(gm < p.M && gk < p.K) va = p.ta ? p.A[(int64_t)gk * p.lda + gm] : p.A[(int64_t)gm * p.lda + gk];
      As[k][m] = va;
      int n;
      if(p.tb) { n = idx / K64; k = idx % K64; } else { k = idx / T64; n = idx % T64; }
      const int gn = n0 + n;
      const int gk2 = k0 + k;
      double vb = 0.0;
      if(gn < p.N && gk2 < p.K) vb = p.tb ? p.B[(int64_t)gn * p.ldb + gk2] : p.B[(int64_t)gk2 * p.ldb + gn];
      Bs[k][n] = vb;
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < K64; ++k) {
      double a[4], b[4];
#pragma unroll
      for(int i = 0; i < 4; ++i) a[i] = As[k][ty + 16 * i];
#pragma unroll
      for(int j = 0; j < 4; ++j) b[j] = Bs[k][tx + 16 * j];
#pragma unroll
      for(int i = 0; i < 4; ++i)
#pragma unroll
        for(int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for(int i = 0; i < 4; ++i)
#pragma unroll
    for(int j = 0; j < 4; ++j) {
      const int gm = m0 + ty + 16 * i, gn = n0 + tx + 16 * j;
      if(gm < p.M && gn < p.N) {
        double* c = p.C + (int64_t)gm * p.ldc + gn;
        double v = p.alpha * acc[i][j];
        if(p.beta != 0.0) v += p.beta * *c;
        *c = v;
      }
    }
}

__global__ void convert_kernel(int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst, int64_t ldd) {
  const int64_t total = rows * cols;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, c = i - r * cols;
    dst[r * ldd + c] = (double)src[r * lds + c];
  }
}

constexpr int RED_BLOCKS = 1024, RED_THREADS = 256;

__device__ inline bool in_part(int uplo, int64_t r, int64_t c) {
  if(uplo == MFX_LOWER) return r >= c;
  if(uplo == MFX_UPPER) return r <= c;
  return true;
}

template <class T>
__device__ inline T block_sum(T v, T* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for(int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if(tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  return sh[0];
}

__global__ __launch_bounds__(RED_THREADS) void frob_partial_kernel(int uplo, int64_t rows, int64_t cols, const float* X,
    int64_t ldx, const double* R, int64_t ldr, double* partial) {
  __shared__ double sh[RED_THREADS];
  const int64_t total = rows * cols;
  double num = 0.0, den = 0.0;
  for(int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < total; i += (int64_t)RED_BLOCKS * RED_THREADS) {
    const int64_t r = i / cols, c = i - r * cols;
    if(!in_part(uplo, r, c)) continue;
    const double ref = R[r * ldr + c];
    const double d = (double)X[r * ldx + c] - ref;
    num = fma(d, d, num);
    den = fma(ref, ref, den);
  }
  const double sn = block_sum(num, sh);
  __syncthreads();
  const double sd = block_sum(den, sh);
  if(threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sn;
    partial[2 * blockIdx.x + 1] = sd;
  }
}

__global__ __launch_bounds__(RED_THREADS) void frob_final_kernel(const double* partial, double* out) {
  __shared__ double sh[RED_THREADS];
  double num = 0.0, den = 0.0;
  for(int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) {
    num += partial[2 * i];
    den += partial[2 * i + 1];
  }
  const double sn = block_sum(num, sh);
  __syncthreads();
  const double sd = block_sum(den, sh);
  if(threadIdx.x == 0) {
    out[0] = sd > 0.0 ? sqrt(sn / sd) : sqrt(sn);
    out[1] = sn;
    out[2] = sd;
  }
}

// percentDiff of polybench/common/polybenchUtilFuncts.h:27-42, evaluated
// with the same float narrowing (absVal takes and returns float).
__device__ inline float abs_val(float a) { return a < 0 ? a * -1 : a; }
__device__ inline float percent_diff(double val1, double val2) {
  if(((double)abs_val((float)val1) < 0.01) && ((double)abs_val((float)val2) < 0.01)) return 0.0f;
  return 100.0f * (abs_val(abs_val((float)(val1 - val2)) / abs_val((float)(val1 + 0.00000001f))));
}

__global__ __launch_bounds__(RED_THREADS) void pdiff_partial_kernel(int64_t rows, int64_t cols, const float* X,
    int64_t ldx, const double* R, int64_t ldr, double threshold, long long* partial) {
  __shared__ long long sh[RED_THREADS];
  const int64_t total = rows * cols;
  long long cnt = 0;
  for(int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < total; i += (int64_t)RED_BLOCKS * RED_THREADS) {
    const int64_t r = i / cols, c = i - r * cols;
    // The reference passes its float result widened to double (verify() loops).
    const double ref = (double)(float)R[r * ldr + c];
    if((double)percent_diff(ref, (double)X[r * ldx + c]) > threshold) ++cnt;
  }
  const long long s = block_sum(cnt, sh);
  if(threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(RED_THREADS) void pdiff_final_kernel(const long long* partial, long long* out) {
  __shared__ long long sh[RED_THREADS];
  long long c = 0;
  for(int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) c += partial[i];
  const long long s = block_sum(c, sh);
  if(threadIdx.x == 0) out[0] = s;
}

__global__ void fill_uniform_kernel(int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed, uint32_t mid) {
  const int64_t total = rows * cols;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, c = i - r * cols;
    p[r * ld + c] = uniform_pm1(keyed_hash(seed, mid, (uint64_t)i));
  }
}

__global__ void fill_lognormal_kernel(
    int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed, uint32_t mid, float mu, float sigma) {
  const int64_t total = rows * cols;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, c = i - r * cols;
    const uint64_t h = keyed_hash(seed, mid, (uint64_t)i);
    // Box-Muller on two 24-bit uniforms in (0, 1); sign from bit 0.
    const float u1 = ((float)(uint32_t)(h >> 40) + 0.5f) * 5.9604644775390625e-08f;
    const float u2 = ((float)(uint32_t)((h >> 16) & 0xffffffu) + 0.5f) * 5.9604644775390625e-08f;
    const float z = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
    const float v = expf(mu + sigma * z);
    p[r * ld + c] = (h & 1u) ? -v : v;
  }
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 16384)); }

}  // namespace

mfx_status launch_gemm_f64(Handle* h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
    int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s) {
  F64Params p{(int)m, (int)n, (int)k, alpha, beta, A, lda, ta, B, ldb, tb, C, ldc};
  dim3 grid((unsigned)ceil_div(n, T64), (unsigned)ceil_div(m, T64));
  if(grid.x == 0 || grid.y == 0) return MFX_SUCCESS;
  return timed_launch(h, "gemm_f64_oracle", s, [&] { gemm_f64_kernel<<<grid, 256, 0, s>>>(p); });
}

mfx_status launch_convert(Handle* h, int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst,
    int64_t ldd, hipStream_t s) {
  if(rows * cols == 0) return MFX_SUCCESS;
  return timed_launch(h, "convert_f32_f64", s,
      [&] { convert_kernel<<<grid_for(rows * cols), 256, 0, s>>>(rows, cols, src, lds, dst, ldd); });
}

mfx_status launch_rel_frobenius(Handle* h, int uplo, int64_t rows, int64_t cols, const float* X, int64_t ldx,
    const double* R, int64_t ldr, double* out_host, hipStream_t s) {
  void* ws = nullptr;
  MFX_CHECK(workspace(h, (2 * RED_BLOCKS + 4) * sizeof(double), &ws));
  double* partial = static_cast<double*>(ws);
  double* result = partial + 2 * RED_BLOCKS;
  MFX_CHECK(timed_launch(h, "rel_frobenius", s, [&] {
    frob_partial_kernel<<<RED_BLOCKS, RED_THREADS, 0, s>>>(uplo, rows, cols, X, ldx, R, ldr, partial);
    frob_final_kernel<<<1, RED_THREADS, 0, s>>>(partial, result);
  }));
  double host[3];
  MFX_HIP(hipMemcpyAsync(host, result, sizeof(host), hipMemcpyDeviceToHost, s));
  MFX_HIP(hipStreamSynchronize(s));
  *out_host = host[0];
  return MFX_SUCCESS;
}

mfx_status launch_percent_diff(Handle* h, int64_t rows, int64_t cols, const float* X, int64_t ldx, const double* R,
    int64_t ldr, double threshold, int64_t* out_host, hipStream_t s) {
  void* ws = nullptr;
  MFX_CHECK(workspace(h, (RED_BLOCKS + 2) * sizeof(long long), &ws));
  long long* partial = static_cast<long long*>(ws);
  long long* result = partial + RED_BLOCKS;
  MFX_CHECK(timed_launch(h, "percent_diff_count", s, [&] {
    pdiff_partial_kernel<<<RED_BLOCKS, RED_THREADS, 0, s>>>(rows, cols, X, ldx, R, ldr, threshold, partial);
    pdiff_final_kernel<<<1, RED_THREADS, 0, s>>>(partial, result);
  }));
  long long host = 0;
  MFX_HIP(hipMemcpyAsync(&host, result, sizeof(host), hipMemcpyDeviceToHost, s));
  MFX_HIP(hipStreamSynchronize(s));
  *out_host = host;
  return MFX_SUCCESS;
}

mfx_status launch_fill_uniform(Handle* h, int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed, uint32_t mid,
    hipStream_t s) {
  if(rows * cols == 0) return MFX_SUCCESS;
  return timed_launch(h, "fill_uniform", s,
      [&] { fill_uniform_kernel<<<grid_for(rows * cols), 256, 0, s>>>(rows, cols, p, ld, seed, mid); });
}

mfx_status launch_fill_lognormal(Handle* h, int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed,
    uint32_t mid, float mu, float sigma, hipStream_t s) {
  if(rows * cols == 0) return MFX_SUCCESS;
  return timed_launch(h, "fill_lognormal", s,
      [&] { fill_lognormal_kernel<<<grid_for(rows * cols), 256, 0, s>>>(rows, cols, p, ld, seed, mid, mu, sigma); });
}

}  // namespace mfx
