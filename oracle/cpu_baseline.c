/*
 * cpu_baseline.c -- OpenMP restatement of the reference's SYCL kernel bodies,
 * standing in for the reference's `--device=cpu` path (no SYCL implementation
 * is installed, so that path cannot be built).  TEST INFRASTRUCTURE ONLY:
 * bench.py times it as the reported CPU baseline; nothing in the library or
 * the benchmark executables links it.
 *
 * Each function keeps the reference kernel's semantics -- one work-item per
 * output element (i, j) with a serial k loop -- mapping work-items onto
 * OpenMP threads (rows i) and SIMD lanes (16 consecutive j), exactly the
 * shape a SYCL CPU device gives an nd-range over range<2>(N, N).  Built with
 * the reference's host release flags (-O3 -march=native -ffast-math,
 * CMakeLists.txt:17) plus -fopenmp.
 */
#include <stddef.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LANES 16

int cpu_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* polybench/gemm.cpp:82-91 over rows [row0, row1):
 * C[i,j] *= beta; for k: C[i,j] += alpha * A[i,k] * B[k,j]. */
void cpu_gemm_rows(const float* A, const float* B, float* C, int64_t n, float alpha, float beta, int64_t row0,
    int64_t row1) {
#pragma omp parallel for schedule(static)
  for(int64_t i = row0; i < row1; ++i) {
    for(int64_t j0 = 0; j0 < n; j0 += LANES) {
      const int64_t w = n - j0 < LANES ? n - j0 : LANES;
      float acc[LANES];
#pragma omp simd
      for(int64_t l = 0; l < LANES; ++l) acc[l] = l < w ? C[i * n + j0 + l] * beta : 0.0f;
      for(int64_t k = 0; k < n; ++k) {
        const float a = alpha * A[i * n + k];
        const float* b = B + k * n + j0;
        if(w == LANES) {
#pragma omp simd
          for(int64_t l = 0; l < LANES; ++l) acc[l] += a * b[l];
        } else {
          for(int64_t l = 0; l < w; ++l) acc[l] += a * b[l];
        }
      }
      for(int64_t l = 0; l < w; ++l) C[i * n + j0 + l] = acc[l];
    }
  }
}

/* The same kernel body on an m x n x k problem (row-major A m x k, B k x n,
 * C m x n), rows [row0, row1): one problem of the batched uneven workload,
 * which the reference would run as one gemm submission per problem. */
void cpu_gemm_mnk_rows(const float* A, const float* B, float* C, int64_t m, int64_t n, int64_t k, float alpha,
    float beta, int64_t row0, int64_t row1) {
  (void)m;
#pragma omp parallel for schedule(static)
  for(int64_t i = row0; i < row1; ++i) {
    for(int64_t j0 = 0; j0 < n; j0 += LANES) {
      const int64_t w = n - j0 < LANES ? n - j0 : LANES;
      float acc[LANES];
#pragma omp simd
      for(int64_t l = 0; l < LANES; ++l) acc[l] = (l < w && beta != 0.0f) ? C[i * n + j0 + l] * beta : 0.0f;
      for(int64_t kk = 0; kk < k; ++kk) {
        const float a = alpha * A[i * k + kk];
        const float* b = B + kk * n + j0;
        if(w == LANES) {
#pragma omp simd
          for(int64_t l = 0; l < LANES; ++l) acc[l] += a * b[l];
        } else {
          for(int64_t l = 0; l < w; ++l) acc[l] += a * b[l];
        }
      }
      for(int64_t l = 0; l < w; ++l) C[i * n + j0 + l] = acc[l];
    }
  }
}

/* polybench/syrk.cpp:74-83 over rows [row0, row1):
 * C[i,j] *= beta; for k: C[i,j] += alpha * A[i,k] * A[j,k]. */
void cpu_syrk_rows(const float* A, float* C, int64_t n, float alpha, float beta, int64_t row0, int64_t row1) {
#pragma omp parallel for schedule(static)
  for(int64_t i = row0; i < row1; ++i) {
    for(int64_t j = 0; j < n; ++j) {
      float c = C[i * n + j] * beta;
#pragma omp simd reduction(+ : c)
      for(int64_t k = 0; k < n; ++k) c += alpha * A[i * n + k] * A[j * n + k];
      C[i * n + j] = c;
    }
  }
}

/* polybench/syr2k.cpp:78-87 over rows [row0, row1). */
void cpu_syr2k_rows(const float* A, const float* B, float* C, int64_t n, float alpha, float beta, int64_t row0,
    int64_t row1) {
#pragma omp parallel for schedule(static)
  for(int64_t i = row0; i < row1; ++i) {
    for(int64_t j = 0; j < n; ++j) {
      float c = C[i * n + j] * beta;
#pragma omp simd reduction(+ : c)
      for(int64_t k = 0; k < n; ++k) c += alpha * A[i * n + k] * B[j * n + k] + alpha * B[i * n + k] * A[j * n + k];
      C[i * n + j] = c;
    }
  }
}

/* runtime/matmulchain.cpp:20-28 (FP32 accumulator) over rows [row0, row1):
 * c = a*b, c written without being read (the reference's discard_write). */
void cpu_matmul_rows(const float* A, const float* B, float* C, int64_t n, int64_t row0, int64_t row1) {
#pragma omp parallel for schedule(static)
  for(int64_t i = row0; i < row1; ++i) {
    for(int64_t j0 = 0; j0 < n; j0 += LANES) {
      const int64_t w = n - j0 < LANES ? n - j0 : LANES;
      float acc[LANES];
#pragma omp simd
      for(int64_t l = 0; l < LANES; ++l) acc[l] = 0.0f;
      for(int64_t k = 0; k < n; ++k) {
        const float a = A[i * n + k];
        const float* b = B + k * n + j0;
        if(w == LANES) {
#pragma omp simd
          for(int64_t l = 0; l < LANES; ++l) acc[l] += a * b[l];
        } else {
          for(int64_t l = 0; l < w; ++l) acc[l] += a * b[l];
        }
      }
      for(int64_t l = 0; l < w; ++l) C[i * n + j0 + l] = acc[l];
    }
  }
}
