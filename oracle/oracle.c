/*
 * oracle.c -- CPU oracle for the dense-contraction path.  TEST INFRASTRUCTURE
 * ONLY: it is loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py to compare against, never by the library or the
 * benchmark executables.
 *
 * Contents:
 *   1. Restatements of the reference benchmarks' input initialisation and
 *      serial CPU loops (the answer each reference verify() compares against),
 *      in plain C with the same float arithmetic, each citing its source.
 *   2. percentDiff, the reference's verification metric.
 *   3. FP64 closed forms of the reference's outputs for its own inputs
 *      (pinned against FP64 brute force in tests/test_oracle.py).
 *   4. FP64 brute-force contraction for arbitrary (seeded) inputs.
 *   5. The keyed SplitMix64 input generator shared bit-for-bit with the GPU.
 *   6. Number formats: RNE BF16 hi/lo split and OCP e4m3 block quantisation
 *      (pinned against exhaustive nearest-value search in the tests).
 *
 * Built strictly (no -ffast-math, -ffp-contract=off) so every result is
 * reproducible; the reference itself builds with -ffast-math
 * (CMakeLists.txt:17), which is why parity with it is measured by its own
 * percentDiff threshold, not bit equality.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef float DATA_TYPE;

/* ------------------------------------------------------------------------ */
/* 1. Reference initialisation and serial loops                              */
/* ------------------------------------------------------------------------ */

/* polybench/gemm.cpp:11-12 -- integer macros promoted to float in use. */
#define GEMM_ALPHA 32412
#define GEMM_BETA 2123

/* polybench/gemm.cpp:18-40 */
void gemm_init(DATA_TYPE* A, DATA_TYPE* B, DATA_TYPE* C, size_t size) {
  const size_t NI = size, NJ = size, NK = size;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NK; j++) A[i * NK + j] = ((DATA_TYPE)i * j) / NI;
  for(size_t i = 0; i < NK; i++)
    for(size_t j = 0; j < NJ; j++) B[i * NJ + j] = ((DATA_TYPE)i * j + 1) / NJ;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NJ; j++) C[i * NJ + j] = ((DATA_TYPE)i * j + 2) / NJ;
}

/* polybench/gemm.cpp:42-56 -- C = BETA*C, then C += ALPHA*A*B term by term. */
void gemm_cpu(const DATA_TYPE* A, const DATA_TYPE* B, DATA_TYPE* C, size_t size) {
  const size_t NI = size, NJ = size, NK = size;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NJ; j++) {
      C[i * NJ + j] *= GEMM_BETA;
      for(size_t k = 0; k < NK; ++k) C[i * NJ + j] += GEMM_ALPHA * A[i * NK + k] * B[k * NJ + j];
    }
}

/* polybench/2mm.cpp:16-45 */
void mm2_init(DATA_TYPE* A, DATA_TYPE* B, DATA_TYPE* C, DATA_TYPE* D, size_t size) {
  const size_t NI = size, NJ = size, NK = size, NL = size;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NK; j++) A[i * NI + j] = ((DATA_TYPE)i * j) / NI;
  for(size_t i = 0; i < NK; i++)
    for(size_t j = 0; j < NJ; j++) B[i * NK + j] = ((DATA_TYPE)i * (j + 1)) / NJ;
  for(size_t i = 0; i < NL; i++)
    for(size_t j = 0; j < NJ; j++) C[i * NL + j] = ((DATA_TYPE)i * (j + 3)) / NL;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NL; j++) D[i * NL + j] = ((DATA_TYPE)i * (j + 2)) / NK;
}

/* polybench/2mm.cpp:47-69 -- C += A*B (C is non-zero on entry), E = C*D. */
void mm2_cpu(const DATA_TYPE* A, const DATA_TYPE* B, DATA_TYPE* C, const DATA_TYPE* D, DATA_TYPE* E, size_t size) {
  const size_t NI = size, NJ = size, NK = size, NL = size;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NJ; j++)
      for(size_t k = 0; k < NK; ++k) C[i * NJ + j] += A[i * NK + k] * B[k * NJ + j];
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NL; j++) {
      E[i * NL + j] = 0;
      for(size_t k = 0; k < NJ; ++k) E[i * NL + j] += C[i * NJ + k] * D[k * NL + j];
    }
}

/* polybench/3mm.cpp:17-47 */
void mm3_init(DATA_TYPE* A, DATA_TYPE* B, DATA_TYPE* C, DATA_TYPE* D, size_t size) {
  const size_t NI = size, NJ = size, NK = size, NL = size, NM = size;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NK; j++) A[i * NK + j] = ((DATA_TYPE)i * j) / NI;
  for(size_t i = 0; i < NK; i++)
    for(size_t j = 0; j < NJ; j++) B[i * NJ + j] = ((DATA_TYPE)i * (j + 1)) / NJ;
  for(size_t i = 0; i < NJ; i++)
    for(size_t j = 0; j < NM; j++) C[i * NM + j] = ((DATA_TYPE)i * (j + 3)) / NL;
  for(size_t i = 0; i < NM; i++)
    for(size_t j = 0; j < NL; j++) D[i * NL + j] = ((DATA_TYPE)i * (j + 2)) / NK;
}

/* polybench/3mm.cpp:49-86 -- E = A*B, F = C*D, G = E*F. */
void mm3_cpu(const DATA_TYPE* A, const DATA_TYPE* B, const DATA_TYPE* C, const DATA_TYPE* D, DATA_TYPE* E, DATA_TYPE* F,
    DATA_TYPE* G, size_t size) {
  const size_t NI = size, NJ = size, NK = size, NL = size, NM = size;
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NJ; j++) {
      E[i * NJ + j] = 0;
      for(size_t k = 0; k < NK; ++k) E[i * NJ + j] += A[i * NK + k] * B[k * NJ + j];
    }
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NL; j++) {
      F[i * NL + j] = 0;
      for(size_t k = 0; k < NM; ++k) F[i * NL + j] += C[i * NM + k] * D[k * NL + j];
    }
  for(size_t i = 0; i < NI; i++)
    for(size_t j = 0; j < NL; j++) {
      G[i * NL + j] = 0;
      for(size_t k = 0; k < NJ; ++k) G[i * NL + j] += E[i * NJ + k] * F[k * NL + j];
    }
}

/* polybench/syrk.cpp:15-16 */
static const DATA_TYPE SYRK_ALPHA = 123;
static const DATA_TYPE SYRK_BETA = 14512;

/* polybench/syrk.cpp:18-31 */
void syrk_init(DATA_TYPE* A, DATA_TYPE* C, size_t size) {
  const size_t N = size, M = size;
  for(size_t i = 0; i < N; i++) {
    for(size_t j = 0; j < M; j++) A[i * M + j] = ((DATA_TYPE)i * j) / N;
    for(size_t j = 0; j < N; j++) C[i * M + j] = ((DATA_TYPE)i * j + 2) / N;
  }
}

/* polybench/syrk.cpp:33-51 -- full square C = beta*C + alpha*A*A^T. */
void syrk_cpu(const DATA_TYPE* A, DATA_TYPE* C, size_t size) {
  const size_t N = size, M = size;
  for(size_t i = 0; i < N; i++)
    for(size_t j = 0; j < N; j++) C[i * M + j] *= SYRK_BETA;
  for(size_t i = 0; i < N; i++)
    for(size_t j = 0; j < N; j++)
      for(size_t k = 0; k < M; k++) C[i * N + j] += SYRK_ALPHA * A[i * M + k] * A[j * M + k];
}

/* polybench/syr2k.cpp:15-16 */
static const DATA_TYPE SYR2K_ALPHA = 1;
static const DATA_TYPE SYR2K_BETA = 1;

/* polybench/syr2k.cpp:18-32 */
void syr2k_init(DATA_TYPE* A, DATA_TYPE* B, DATA_TYPE* C, size_t size) {
  const size_t N = size, M = size;
  for(size_t i = 0; i < N; i++) {
    for(size_t j = 0; j < N; j++) C[i * N + j] = ((DATA_TYPE)i * j + 2) / N;
    for(size_t j = 0; j < M; j++) {
      A[i * N + j] = ((DATA_TYPE)i * j) / N;
      B[i * N + j] = ((DATA_TYPE)i * j + 1) / N;
    }
  }
}

/* polybench/syr2k.cpp:34-52 -- two separate += per k, as the CPU loop does. */
void syr2k_cpu(const DATA_TYPE* A, const DATA_TYPE* B, DATA_TYPE* C, size_t size) {
  const size_t N = size, M = size;
  for(size_t i = 0; i < N; i++)
    for(size_t j = 0; j < N; j++) C[i * N + j] *= SYR2K_BETA;
  for(size_t i = 0; i < N; i++)
    for(size_t j = 0; j < N; j++)
      for(size_t k = 0; k < M; k++) {
        C[i * N + j] += SYR2K_ALPHA * A[i * M + k] * B[j * M + k];
        C[i * N + j] += SYR2K_ALPHA * B[i * M + k] * A[j * M + k];
      }
}

/* polybench/covariance.cpp:18 */
static const DATA_TYPE COV_FLOAT_N = 3214212.01f;

/* polybench/covariance.cpp:20-29 -- into a zeroed (size+1)^2 array, stride
 * size+1: rows and columns 0..size-1 set, row and column `size` left zero. */
void covariance_init(DATA_TYPE* data, size_t size) {
  const size_t M = size, N = size;
  for(size_t i = 0; i < M; i++)
    for(size_t j = 0; j < N; j++) data[i * (N + 1) + j] = ((DATA_TYPE)i * j) / M;
}

/* polybench/covariance.cpp:31-61 -- 1-based; `data` is centred in place. */
void covariance_cpu(DATA_TYPE* data, DATA_TYPE* symmat, DATA_TYPE* mean, size_t size) {
  const size_t M = size, N = size;
  for(size_t j = 1; j <= M; j++) {
    mean[j] = 0.0f;
    for(size_t i = 1; i <= N; i++) mean[j] += data[i * (M + 1) + j];
    mean[j] /= COV_FLOAT_N;
  }
  for(size_t i = 1; i <= N; i++)
    for(size_t j = 1; j <= M; j++) data[i * (M + 1) + j] -= mean[j];
  for(size_t j1 = 1; j1 <= M; j1++)
    for(size_t j2 = j1; j2 <= M; j2++) {
      symmat[j1 * (M + 1) + j2] = 0.0f;
      for(size_t i = 1; i <= N; i++) symmat[j1 * (M + 1) + j2] += data[i * (M + 1) + j1] * data[i * (M + 1) + j2];
      symmat[j2 * (M + 1) + j1] = symmat[j1 * (M + 1) + j2];
    }
}

/* polybench/correlation.cpp:8-9 */
static const double CORR_FLOAT_N = 3214212.01;
static const double CORR_EPS = 0.005;

/* polybench/correlation.cpp:17-25 -- into a zeroed (size+1)^2 array with row
 * stride `size` (not size+1), i and j from 0 to size inclusive; later writes
 * win where flat indices coincide. */
void correlation_init(DATA_TYPE* data, size_t size) {
  const size_t M = size, N = size;
  for(size_t i = 0; i <= M; i++)
    for(size_t j = 0; j <= N; j++) data[i * N + j] = ((DATA_TYPE)i * j) / (M + 1);
}

/* polybench/correlation.cpp:27-70 -- 1-based; float statistics divided by the
 * double constant where the reference divides a float by a double. */
void correlation_cpu(DATA_TYPE* data, DATA_TYPE* mean, DATA_TYPE* stddev, DATA_TYPE* symmat, size_t size) {
  const size_t M = size, N = size;
  for(size_t j = 1; j <= M; j++) {
    mean[j] = 0.0f;
    for(size_t i = 1; i <= N; i++) mean[j] += data[i * (M + 1) + j];
    mean[j] /= (DATA_TYPE)CORR_FLOAT_N;
  }
  for(size_t j = 1; j <= M; j++) {
    stddev[j] = 0.0f;
    for(size_t i = 1; i <= N; i++)
      stddev[j] += (data[i * (M + 1) + j] - mean[j]) * (data[i * (M + 1) + j] - mean[j]);
    stddev[j] = (DATA_TYPE)((double)stddev[j] / CORR_FLOAT_N);
    stddev[j] = (DATA_TYPE)sqrt((double)stddev[j]);
    stddev[j] = stddev[j] <= CORR_EPS ? 1.0f : stddev[j];
  }
  for(size_t i = 1; i <= N; i++)
    for(size_t j = 1; j <= M; j++) {
      data[i * (M + 1) + j] -= mean[j];
      data[i * (M + 1) + j] = (DATA_TYPE)((double)data[i * (M + 1) + j] / sqrt(CORR_FLOAT_N));
      data[i * (M + 1) + j] /= stddev[j];
    }
  for(size_t j1 = 1; j1 <= M - 1; j1++) {
    symmat[j1 * (M + 1) + j1] = 1.0f;
    for(size_t j2 = j1 + 1; j2 <= M; j2++) {
      symmat[j1 * (M + 1) + j2] = 0.0f;
      for(size_t i = 1; i <= N; i++) symmat[j1 * (M + 1) + j2] += data[i * (M + 1) + j1] * data[i * (M + 1) + j2];
      symmat[j2 * (M + 1) + j1] = symmat[j1 * (M + 1) + j2];
    }
  }
  symmat[M * (M + 1) + M] = 1.0f;
}

/* runtime/matmulchain.cpp:62-70 -- identity inputs. */
void matmulchain_identity(DATA_TYPE* M, size_t n) {
  for(size_t i = 0; i < n; ++i)
    for(size_t j = 0; j < n; ++j) M[i * n + j] = i == j;
}

/* runtime/matmulchain.cpp:20-28 -- `auto sum = 0` makes the accumulator an
 * int: every step computes (float)sum + a*b and truncates back to int. */
void matmulchain_multiply_int(const DATA_TYPE* a, const DATA_TYPE* b, DATA_TYPE* c, size_t n) {
  for(size_t i = 0; i < n; ++i)
    for(size_t j = 0; j < n; ++j) {
      int sum = 0;
      for(size_t k = 0; k < n; ++k) {
        const DATA_TYPE a_ik = a[i * n + k];
        const DATA_TYPE b_kj = b[k * n + j];
        sum += a_ik * b_kj;
      }
      c[i * n + j] = sum;
    }
}

/* ------------------------------------------------------------------------ */
/* 2. percentDiff -- polybench/common/polybenchUtilFuncts.h:14,27-42         */
/* ------------------------------------------------------------------------ */
#define SMALL_FLOAT_VAL 0.00000001f

static float absVal(float a) { return a < 0 ? (a * -1) : a; }

float percent_diff(double val1, double val2) {
  if((absVal(val1) < 0.01) && (absVal(val2) < 0.01)) return 0.0f;
  return 100.0f * (absVal(absVal(val1 - val2) / absVal(val1 + SMALL_FLOAT_VAL)));
}

/* Number of elements failing `percentDiff(ref, x) > threshold` -- the loop of
 * every polybench verify(), e.g. polybench/gemm.cpp:107-113. */
int64_t percent_diff_count(const DATA_TYPE* ref, const DATA_TYPE* x, size_t count, double threshold) {
  int64_t bad = 0;
  for(size_t i = 0; i < count; ++i)
    if(percent_diff(ref[i], x[i]) > threshold) ++bad;
  return bad;
}

/* ------------------------------------------------------------------------ */
/* 3. FP64 closed forms of the reference outputs for the reference inputs.   */
/*    S1 = N(N-1)/2, S2 = (N-1)N(2N-1)/6 (sums of k and k^2 over k < N).     */
/* ------------------------------------------------------------------------ */
static double s1(double n) { return n * (n - 1) / 2; }
static double s2(double n) { return (n - 1) * n * (2 * n - 1) / 6; }

/* gemm: C_ij = 2123 (ij+2)/N + 32412 (i/N^2)(j S2 + S1). */
void gemm_closed(double* out, size_t size) {
  const double N = (double)size, S1 = s1(N), S2 = s2(N);
  for(size_t i = 0; i < size; ++i)
    for(size_t j = 0; j < size; ++j)
      out[i * size + j] = GEMM_BETA * ((double)i * j + 2) / N + GEMM_ALPHA * ((double)i / (N * N)) * ((double)j * S2 + S1);
}

/* 2mm: E_ij = (i(j+2)/N) [(S2 + 3 S1)/N + S2 (S2 + S1)/N^2]. */
void mm2_closed(double* out, size_t size) {
  const double N = (double)size, S1 = s1(N), S2 = s2(N);
  const double f = (S2 + 3 * S1) / N + S2 * (S2 + S1) / (N * N);
  for(size_t i = 0; i < size; ++i)
    for(size_t j = 0; j < size; ++j) out[i * size + j] = ((double)i * (j + 2) / N) * f;
}

/* 3mm: G_ij = i (j+2) S2 (S2 + 3 S1)(S2 + S1) / N^4. */
void mm3_closed(double* out, size_t size) {
  const double N = (double)size, S1 = s1(N), S2 = s2(N);
  const double f = S2 * (S2 + 3 * S1) * (S2 + S1) / (N * N * N * N);
  for(size_t i = 0; i < size; ++i)
    for(size_t j = 0; j < size; ++j) out[i * size + j] = (double)i * (j + 2) * f;
}

/* syrk: C_ij = 14512 (ij+2)/N + 123 ij S2 / N^2. */
void syrk_closed(double* out, size_t size) {
  const double N = (double)size, S2 = s2(N);
  for(size_t i = 0; i < size; ++i)
    for(size_t j = 0; j < size; ++j)
      out[i * size + j] = 14512.0 * ((double)i * j + 2) / N + 123.0 * (double)i * j * S2 / (N * N);
}

/* syr2k: C_ij = (ij+2)/N + (2 ij S2 + (i+j) S1) / N^2. */
void syr2k_closed(double* out, size_t size) {
  const double N = (double)size, S1 = s1(N), S2 = s2(N);
  for(size_t i = 0; i < size; ++i)
    for(size_t j = 0; j < size; ++j)
      out[i * size + j] = ((double)i * j + 2) / N + (2.0 * i * j * S2 + ((double)i + j) * S1) / (N * N);
}

/* ------------------------------------------------------------------------ */
/* 4. FP64 brute-force contraction                                           */
/* ------------------------------------------------------------------------ */
/* C = alpha op(A) op(B) + beta C (beta == 0: C not read), sequential k. */
void gemm_f64(int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
    const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  double* row = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for(int64_t i = 0; i < m; ++i) {
    for(int64_t j = 0; j < n; ++j) row[j] = 0.0;
    for(int64_t p = 0; p < k; ++p) {
      const double a = ta ? A[p * lda + i] : A[i * lda + p];
      if(tb)
        for(int64_t j = 0; j < n; ++j) row[j] += a * B[j * ldb + p];
      else
        for(int64_t j = 0; j < n; ++j) row[j] += a * B[p * ldb + j];
    }
    for(int64_t j = 0; j < n; ++j) C[i * ldc + j] = alpha * row[j] + (beta != 0.0 ? beta * C[i * ldc + j] : 0.0);
  }
  free(row);
}

/* ------------------------------------------------------------------------ */
/* 5. Keyed SplitMix64 inputs (bit-identical to the device generator)        */
/* ------------------------------------------------------------------------ */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* The SplitMix64 output function, exported so tests can pin it against the
 * published SplitMix64 sequence. */
uint64_t splitmix64_mix(uint64_t z) { return mix64(z); }

uint64_t keyed_hash(uint64_t seed, uint32_t matrix, uint64_t index) {
  const uint64_t base = mix64(seed ^ (0x9e3779b97f4a7c15ull * ((uint64_t)matrix + 1ull)));
  return mix64(base + 0x9e3779b97f4a7c15ull * (index + 1ull));
}

/* p[r*ld + c] = 2 (h >> 40) 2^-24 - 1 with h keyed by (seed, matrix, r*cols + c). */
void fill_uniform(float* p, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, uint32_t matrix) {
  for(int64_t r = 0; r < rows; ++r)
    for(int64_t c = 0; c < cols; ++c) {
      const uint64_t h = keyed_hash(seed, matrix, (uint64_t)(r * cols + c));
      p[r * ld + c] = (float)(uint32_t)(h >> 40) * 1.1920928955078125e-07f - 1.0f;
    }
}

/* sign * exp(mu + sigma * z): z from Box-Muller on two 24-bit uniforms of
 * the same keyed hash, u1 = (bits 40..63 + 1/2) 2^-24, u2 = (bits 16..39 +
 * 1/2) 2^-24, each formed in float (so the + 1/2 rounds away for values of
 * 2^23 and above, and u1 can be exactly 1), z = sqrt(-2 ln u1) cos(2 pi u2);
 * negative when bit 0 is set.  The uniforms are bit-exact with the library;
 * the rest is evaluated in double here and in float there, so the two agree
 * to a few float ulps, not bit for bit.  No reference counterpart: the
 * block-scaled FP8 workload's input generator. */
void fill_lognormal(float* p, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, uint32_t matrix, double mu,
    double sigma) {
  const double two_pi = 6.283185307179586476925286766559;
  for(int64_t r = 0; r < rows; ++r)
    for(int64_t c = 0; c < cols; ++c) {
      const uint64_t h = keyed_hash(seed, matrix, (uint64_t)(r * cols + c));
      const float f1 = ((float)(uint32_t)(h >> 40) + 0.5f) * 5.9604644775390625e-08f;
      const float f2 = ((float)(uint32_t)((h >> 16) & 0xffffffu) + 0.5f) * 5.9604644775390625e-08f;
      const double u1 = f1, u2 = f2;
      const double z = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
      const double v = exp(mu + sigma * z);
      p[r * ld + c] = (float)((h & 1u) ? -v : v);
    }
}

/* ------------------------------------------------------------------------ */
/* 6. Number formats                                                         */
/* ------------------------------------------------------------------------ */
static uint32_t fbits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}
static float bitsf(uint32_t u) {
  float x;
  memcpy(&x, &u, 4);
  return x;
}

/* FP32 -> BF16 with round-to-nearest-even (NaN kept quiet). */
uint16_t bf16_rne(float x) {
  uint32_t u = fbits(x);
  if((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_to_f32(uint16_t h) { return bitsf((uint32_t)h << 16); }

/* x = hi + lo + r: hi = bf16(x), lo = bf16(x - hi). */
void split_bf16(const float* x, uint16_t* hi, uint16_t* lo, int64_t n) {
  for(int64_t i = 0; i < n; ++i) {
    hi[i] = bf16_rne(x[i]);
    lo[i] = bf16_rne(x[i] - bf16_to_f32(hi[i]));
  }
}

/* FP32 -> OCP e4m3 (E4M3FN), round-to-nearest-even, saturating to +-448. */
uint8_t e4m3_rne_sat(float x) {
  const uint32_t u = fbits(x);
  const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
  const uint32_t au = u & 0x7fffffffu;
  if(au > 0x7f800000u) return (uint8_t)(sign | 0x7f);
  const float a = bitsf(au);
  if(a > 448.0f) return (uint8_t)(sign | 0x7e);
  if(a < 0.015625f) {
    const float s = a * 512.0f;
    uint32_t i = (uint32_t)s;
    const float frac = s - (float)i;
    if(frac > 0.5f || (frac == 0.5f && (i & 1u))) ++i;
    return (uint8_t)(sign | i);
  }
  const int e = (int)((au >> 23) & 0xffu) - 127;
  const uint32_t m = au & 0x7fffffu;
  const uint32_t r = (m + 0x7ffffu + ((m >> 20) & 1u)) >> 20;
  uint32_t code = ((uint32_t)(e + 7) << 3) + r;
  if(code > 0x7eu) code = 0x7eu;
  return (uint8_t)(sign | code);
}

float e4m3_to_f32(uint8_t q) {
  const uint32_t s = q & 0x80u, e = (q >> 3) & 0xfu, m = q & 7u;
  float v;
  if(e == 15 && m == 7) return bitsf(0x7fc00000u | (s << 24));
  if(e == 0)
    v = (float)m * 0.001953125f;
  else
    v = bitsf(((e + 120u) << 23) | (m << 20));
  return s ? -v : v;
}

/* Smallest e with amax <= 448 * 2^e (0 when amax == 0), clamped to [-127, 127]. */
int fp8_block_exponent(float amax) {
  if(!(amax > 0.0f)) return 0;
  int e;
  if(((fbits(amax) >> 23) & 0xffu) == 0)
    e = -134;
  else
    e = (int)((fbits(amax) >> 23) & 0xffu) - 127 - 8;
  for(int it = 0; it < 4; ++it) {
    if(amax > ldexpf(448.0f, e))
      ++e;
    else if(amax <= ldexpf(448.0f, e - 1))
      --e;
    else
      break;
  }
  if(e < -127) e = -127;
  if(e > 127) e = 127;
  return e;
}

/* Per block x block tile: e = fp8_block_exponent(max |x|), q = e4m3(x * 2^-e). */
void quantize_fp8_block(const float* x, int64_t rows, int64_t cols, int64_t ldx, int block, uint8_t* q, int32_t* exps) {
  const int64_t tr = (rows + block - 1) / block, tc = (cols + block - 1) / block;
  for(int64_t bi = 0; bi < tr; ++bi)
    for(int64_t bj = 0; bj < tc; ++bj) {
      float amax = 0.0f;
      for(int64_t r = bi * block; r < rows && r < (bi + 1) * block; ++r)
        for(int64_t c = bj * block; c < cols && c < (bj + 1) * block; ++c) {
          const float a = fabsf(x[r * ldx + c]);
          if(a > amax) amax = a;
        }
      const int e = fp8_block_exponent(amax);
      exps[bi * tc + bj] = e;
      const float inv = ldexpf(1.0f, -e);
      for(int64_t r = bi * block; r < rows && r < (bi + 1) * block; ++r)
        for(int64_t c = bj * block; c < cols && c < (bj + 1) * block; ++c) q[r * cols + c] = e4m3_rne_sat(x[r * ldx + c] * inv);
    }
}
