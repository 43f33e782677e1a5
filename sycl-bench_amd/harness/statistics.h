// Shared pieces of the covariance / correlation benchmarks: the reference's
// (N+1) x (N+1) 1-based storage, and an FP64 check of the m x m result.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

#include "contraction.h"

// The interior of a 1-based (n1 x n1) array: element (1, 1) onward, ld = n1.
template <class T>
inline T* interior(DeviceBuffer<T>& b, std::size_t n1) {
  return b.get() + n1 + 1;
}

// FP64 covariance (correlation: also scaled, with a unit diagonal) of the
// n x m row-major host matrix x (leading dimension ld) as an m x m device
// matrix: statistics on the host in double, the m x m x n product on the GPU.
inline DeviceBuffer<double> statsF64(HipDevice& d, const std::vector<float>& x, std::size_t n, std::size_t m,
    std::size_t ld, double n_const, bool correlation, double eps) {
  std::vector<double> mean(m, 0.0), sd(m, 1.0);
  for(std::size_t i = 0; i < n; ++i)
    for(std::size_t j = 0; j < m; ++j) mean[j] += x[i * ld + j];
  for(auto& v : mean) v /= n_const;
  if(correlation) {
    std::vector<double> s(m, 0.0);
    for(std::size_t i = 0; i < n; ++i)
      for(std::size_t j = 0; j < m; ++j) {
        const double c = x[i * ld + j] - mean[j];
        s[j] += c * c;
      }
    for(std::size_t j = 0; j < m; ++j) {
      const double v = std::sqrt(s[j] / n_const);
      sd[j] = v <= eps ? 1.0 : v;
    }
  }
  const double scale = correlation ? std::sqrt(n_const) : 1.0;
  std::vector<double> D(n * m);
  for(std::size_t i = 0; i < n; ++i)
    for(std::size_t j = 0; j < m; ++j) D[i * m + j] = (x[i * ld + j] - mean[j]) / scale / sd[j];
  DeviceBuffer<double> Dd, R;
  Dd.initialize(d, D.data(), D.size());
  R.initialize(d, m * m);
  mfxCheck(mfx_gemm_f64(d.handle(), MFX_OP_T, MFX_OP_N, m, m, n, 1.0, Dd.get(), m, Dd.get(), m, 0.0, R.get(), m,
               d.stream()),
      "mfx_gemm_f64");
  if(correlation) {
    std::vector<double> r = R.toHost(d);
    for(std::size_t j = 0; j < m; ++j) r[j * m + j] = 1.0;
    R.initialize(d, r.data(), r.size());
  }
  d.synchronize();
  return R;
}

// Dense m x m copy of the interior of a 1-based n1 x n1 float matrix.
inline DeviceBuffer<float> denseInterior(HipDevice& d, DeviceBuffer<float>& b, std::size_t n1, std::size_t m) {
  DeviceBuffer<float> out;
  out.initialize(d, m * m);
  hipCheck(hipMemcpy2DAsync(out.get(), m * sizeof(float), interior(b, n1), n1 * sizeof(float), m * sizeof(float), m,
               hipMemcpyDeviceToDevice, d.stream()),
      "hipMemcpy2DAsync");
  return out;
}

// Host copy of the n x m interior of a 1-based n1 x n1 float matrix (dense).
inline std::vector<float> hostInterior(HipDevice& d, DeviceBuffer<float>& b, std::size_t n1, std::size_t n,
    std::size_t m) {
  std::vector<float> h(n * m);
  hipCheck(hipMemcpy2DAsync(h.data(), m * sizeof(float), interior(b, n1), n1 * sizeof(float), m * sizeof(float), n,
               hipMemcpyDeviceToHost, d.stream()),
      "hipMemcpy2DAsync");
  d.synchronize();
  return h;
}
