/*
 * mfx.h - C ABI of the MI355X (gfx950) dense-contraction library behind the
 * polybench gemm / 2mm / 3mm / syrk / syr2k and runtime/matmulchain benchmarks.
 *
 * The reference suite (SYCL-Bench) has no library boundary: each benchmark's
 * run() submits a SYCL parallel_for.  Every entry point below replaces one of
 * those kernel submissions (cited per function), so the benchmark classes call
 * only this header for GPU work.
 *
 * Conventions (all entry points):
 *   - Matrices are row-major FP32 unless stated; element (r, c) of a matrix
 *     with leading dimension ld sits at p[r * ld + c].
 *   - Pointers are caller-owned device pointers.  The library owns only a
 *     per-handle workspace (split BF16 / FP8 panels, scales, partial sums).
 *   - All work is enqueued on `stream` (a hipStream_t passed as void*; NULL is
 *     the null stream) and is asynchronous unless a function says "blocking".
 *   - Functions return an mfx_status; nothing throws or aborts across the ABI.
 *   - A handle must not be used from two host threads at once; distinct
 *     handles are independent.
 */
#ifndef MFX_H
#define MFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mfx_status {
  MFX_SUCCESS = 0,
  MFX_ERROR_INVALID_VALUE = 1,   /* bad size, pointer, enum or leading dimension */
  MFX_ERROR_NOT_SUPPORTED = 2,   /* valid request that the library does not implement */
  MFX_ERROR_HIP = 3,             /* HIP runtime error (see mfx_last_error) */
  MFX_ERROR_ALLOC = 4,           /* workspace allocation failed */
  MFX_ERROR_NO_DEVICE = 5,       /* no usable gfx950 device */
  MFX_ERROR_TIMEOUT = 6,         /* a bounded in-kernel wait expired (grouped launch) */
  MFX_ERROR_IO = 7               /* autotune cache file could not be read or written */
} mfx_status;

typedef enum mfx_precision {
  MFX_PREC_FP32 = 0,        /* FP32 operands on v_mfma_f32_32x32x2_f32, FP32 accumulate */
  MFX_PREC_SPLIT_BF16 = 1,  /* x = hi + lo (both BF16, RNE); hi*hi + hi*lo + lo*hi on BF16 MFMA, FP32 accumulate */
  MFX_PREC_BF16 = 2,        /* single RNE BF16 rounding of each operand, FP32 accumulate */
  MFX_PREC_FP8_BLOCK = 3,   /* OCP e4m3 with one power-of-two scale per 128x128 block, FP32 accumulate */
  MFX_PREC_FP8_BLOCK2 = 4   /* two-term e4m3: x ~ s1*q1 + s2*q2; q1*q1 + q1*q2 + q2*q1 */
} mfx_precision;

typedef enum mfx_op { MFX_OP_N = 0, MFX_OP_T = 1 } mfx_op;
typedef enum mfx_uplo { MFX_LOWER = 0, MFX_UPPER = 1, MFX_FULL = 2 } mfx_uplo;

typedef struct mfx_handle_s* mfx_handle;
typedef void* mfx_stream; /* hipStream_t */

/* ---- handle ------------------------------------------------------------ */

/* Creates a handle bound to HIP device `device` (which becomes current). */
mfx_status mfx_create(mfx_handle* handle, int device);
mfx_status mfx_destroy(mfx_handle handle);
/* Pre-allocates `bytes` of workspace so later calls (and graph capture) do not allocate. */
mfx_status mfx_reserve_workspace(mfx_handle handle, size_t bytes);
/* Workspace bytes a gemm of this shape and precision needs under the
 * automatic choices on the current device (FP32: split-K partial tiles of a
 * small grid, at most 32 MiB, else 0). */
size_t mfx_gemm_workspace_size(mfx_precision prec, int64_t m, int64_t n, int64_t k);
/*
 * Host-only pre-flight check (no handle, no device): the status mfx_gemm
 * returns for these enums, shapes and leading dimensions before it touches
 * any memory -- MFX_ERROR_INVALID_VALUE, MFX_ERROR_NOT_SUPPORTED (K beyond
 * 2^25 for SPLIT_BF16 or 2^26 for BF16, the packed paths' 32-bit panel
 * offsets) or MFX_SUCCESS.  Operand pointers are not checked.
 * No reference counterpart (SYCL-Bench validates nothing before submitting).
 */
mfx_status mfx_gemm_check(mfx_precision prec, mfx_op transA, mfx_op transB, int64_t m, int64_t n, int64_t k,
    int64_t lda, int64_t ldb, int64_t ldc);
const char* mfx_status_string(mfx_status status);
/* Text of the last error seen on the calling thread (HIP error names included). */
const char* mfx_last_error(void);
/* Device name as reported by HIP (e.g. "AMD Instinct MI355X"), blocking. */
mfx_status mfx_device_name(mfx_handle handle, char* buf, size_t len);

/* ---- dense contractions -------------------------------------------------- */

/* Size of the fixed device array behind mfx_gemm_noalloc (one per device). */
#define MFX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/*
 * C = alpha * op(A) * op(B) + beta * C, op(A) is m x k, op(B) is k x n.
 * beta == 0 means C is not read (BLAS convention).
 * Replaces: polybench/gemm.cpp:74-93 (Gemm kernel, C = BETA*C + ALPHA*A*B),
 *           polybench/2mm.cpp:94-123 (Polybench_2mm_1 with alpha=beta=1,
 *           Polybench_2mm_2 with beta=0), polybench/3mm.cpp:115-158
 *           (Polybench_3mm_1/_2/_3 into zeroed outputs, beta=0),
 *           runtime/matmulchain.cpp:12-30 (multiply<T>, beta=0).
 */
mfx_status mfx_gemm(mfx_handle handle, mfx_precision prec, mfx_op transA, mfx_op transB,
    int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
    const float* B, int64_t ldb, float beta, float* C, int64_t ldc, mfx_stream stream);

/*
 * mfx_gemm without any allocation: the same kernels and results as mfx_gemm
 * (FP32, SPLIT_BF16 or BF16), with the packed operands (or FP32 split-K
 * partial tiles, at most 32 MiB) in a fixed library array of
 * MFX_NOALLOC_SCRATCH_BYTES per device instead of the handle's workspace.  It calls no hipMalloc / hipFree or other allocating routine,
 * so code under graph capture, tight serving loops and harnesses that
 * forbid allocation can call it.  A request whose packed operands exceed the
 * array (4 (m + n) k bytes for SPLIT_BF16, half that for BF16, about) returns
 * MFX_ERROR_NOT_SUPPORTED; FP8 precisions are not supported.  Every call on
 * one device shares that device's array, so calls that overlap on one
 * device race.  Its kernels are not recorded by mfx_timing_*.
 * No reference counterpart: it is for callers that cannot allocate.
 */
mfx_status mfx_gemm_noalloc(mfx_handle handle, mfx_precision prec, mfx_op transA, mfx_op transB,
    int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
    const float* B, int64_t ldb, float beta, float* C, int64_t ldc, mfx_stream stream);

/*
 * C = alpha * A * A^T + beta * C on the `uplo` triangle of the n x n matrix C
 * (A is n x k).  Only tiles of that triangle are scheduled.  With mirror != 0
 * the opposite triangle is written too, each element computed from its own C
 * value (beta * C[j][i] + alpha * (A A^T)[j][i]), which reproduces the
 * reference's full-square result.
 * Replaces: polybench/syrk.cpp:67-85 (kernel Syr2k2, full square).
 */
mfx_status mfx_syrk(mfx_handle handle, mfx_precision prec, mfx_uplo uplo, int mirror,
    int64_t n, int64_t k, float alpha, const float* A, int64_t lda, float beta, float* C,
    int64_t ldc, mfx_stream stream);

/*
 * C = alpha * (A * B^T + B * A^T) + beta * C on the `uplo` triangle (A, B n x k),
 * mirror as in mfx_syrk.
 * Replaces: polybench/syr2k.cpp:70-89 (kernel Syr2k1, full square).
 */
mfx_status mfx_syr2k(mfx_handle handle, mfx_precision prec, mfx_uplo uplo, int mirror,
    int64_t n, int64_t k, float alpha, const float* A, int64_t lda, const float* B,
    int64_t ldb, float beta, float* C, int64_t ldc, mfx_stream stream);

/* ---- covariance / correlation --------------------------------------------- */

/*
 * Covariance of the n x m row-major `data` (n observations of m variables):
 * mean[j] = (sum_i data[i][j]) / n_const (float division), centered = data -
 * mean (n x m, ldw), symmat = centeredᵀ · centered (m x m, full square: upper
 * triangle computed, mirrored).  `prec` selects the contraction (FP32,
 * SPLIT_BF16, BF16).  `data` is not modified, so repeated calls agree.
 * Replaces: polybench/covariance.cpp:76-128 (kernels CovarianceMean,
 * CovarianceReduce, CovarianceCovar; the reference centres `data` in place).
 */
mfx_status mfx_covariance(mfx_handle handle, mfx_precision prec, int64_t m, int64_t n,
    const float* data, int64_t ld, float n_const, float* mean, float* centered, int64_t ldw,
    float* symmat, int64_t lds, mfx_stream stream);
/*
 * Correlation of the n x m `data`: mean as above with (float)n_const;
 * stddev[j] = sqrt((sum_i (data[i][j] - mean[j])^2) / n_const), replaced by 1
 * when <= eps; centered = (data - mean) / sqrt(n_const) / stddev; symmat =
 * centeredᵀ · centered (full square) with a unit diagonal.
 * Replaces: polybench/correlation.cpp:92-158 (kernels CorrelationMean,
 * CorrelationStd, CorrelationReduce, CorrelationCorr, Correlation5; the
 * reference accumulates mean / stddev onto their previous contents and
 * centres `data` in place, so its repeated runs drift; these calls do not).
 */
mfx_status mfx_correlation(mfx_handle handle, mfx_precision prec, int64_t m, int64_t n,
    const float* data, int64_t ld, double n_const, double eps, float* mean, float* stddev,
    float* centered, int64_t ldw, float* symmat, int64_t lds, mfx_stream stream);

/* ---- grouped / chained launches ------------------------------------------ */

typedef struct mfx_gemm_problem {
  int64_t m, n, k;
  float alpha, beta;
  const float* A; int64_t lda; int32_t transA;
  const float* B; int64_t ldb; int32_t transB;
  float* C; int64_t ldc;
} mfx_gemm_problem;

/* Problem `consumer` reads the C of problem `producer` as an operand. */
typedef struct mfx_dep {
  int32_t consumer;
  int32_t producer;
} mfx_dep;

/*
 * Runs `count` problems in one persistent grouped launch.  Producers listed in
 * `deps` must come before their consumers in `problems`; a consumer tile waits
 * in-kernel (bounded) for its producers.  Precisions: FP32, SPLIT_BF16, BF16.
 * No reference counterpart: it batches the per-kernel submissions of
 * polybench/2mm.cpp:91-124, polybench/3mm.cpp:112-159 and
 * runtime/matmulchain.cpp:81-85 (whose DAG edges come from SYCL accessors).
 */
mfx_status mfx_gemm_grouped(mfx_handle handle, mfx_precision prec,
    const mfx_gemm_problem* problems, int32_t count, const mfx_dep* deps, int32_t ndeps,
    mfx_stream stream);

/*
 * Product of `count` n x n matrices mats[0] * mats[1] * ... in balanced-tree
 * order (((M0 M1)(M2 M3))((M4 M5)(M6 M7)) for 8), as one grouped launch.
 * `scratch` holds count-2 device buffers of n*n floats for intermediates.
 * Generalises runtime/matmulchain.cpp:81-85 ((A B)(C D)).
 */
mfx_status mfx_chain(mfx_handle handle, mfx_precision prec, int32_t count,
    const float* const* mats, int64_t n, float* const* scratch, float* out, mfx_stream stream);

/* Blocking: waits for `stream`, then reports MFX_ERROR_TIMEOUT if an in-kernel
 * dependency wait of the handle's last mfx_gemm_grouped / mfx_chain launch
 * exceeded its spin limit (the launch then skipped work; outputs are invalid). */
mfx_status mfx_grouped_status(mfx_handle handle, mfx_stream stream);

/* Graph form of mfx_gemm_grouped: captures the grouped launch into a hipGraph
 * once (instantiate) and replays it (launch).  The problem list is copied; the
 * graph owns its work tables and counters (device memory freed by
 * mfx_graph_destroy).  Replays of one graph must not overlap. */
typedef struct mfx_graph_s* mfx_graph;
mfx_status mfx_grouped_graph_create(mfx_handle handle, mfx_precision prec,
    const mfx_gemm_problem* problems, int32_t count, const mfx_dep* deps, int32_t ndeps,
    mfx_stream stream, mfx_graph* graph);
mfx_status mfx_graph_launch(mfx_graph graph, mfx_stream stream);
/* Blocking: like mfx_grouped_status for the graph's last replay. */
mfx_status mfx_graph_status(mfx_graph graph, mfx_stream stream);
mfx_status mfx_graph_destroy(mfx_graph graph);

/* ---- FP64 oracle and error metrics (scoring only, not part of any product) ---- */

/* C = alpha * op(A) * op(B) + beta * C entirely in FP64 (beta == 0: C not read). */
mfx_status mfx_gemm_f64(mfx_handle handle, mfx_op transA, mfx_op transB, int64_t m,
    int64_t n, int64_t k, double alpha, const double* A, int64_t lda, const double* B,
    int64_t ldb, double beta, double* C, int64_t ldc, mfx_stream stream);
/* dst[r*ldd + c] = (double)src[r*lds + c] */
mfx_status mfx_convert_f32_f64(mfx_handle handle, int64_t rows, int64_t cols,
    const float* src, int64_t lds, double* dst, int64_t ldd, mfx_stream stream);
/*
 * ||X - R||_F / ||R||_F over the `uplo` part of a rows x cols matrix, summed in
 * FP64 in a fixed order (deterministic).  Blocking: result written to *out.
 */
mfx_status mfx_rel_frobenius(mfx_handle handle, mfx_uplo uplo, int64_t rows, int64_t cols,
    const float* X, int64_t ldx, const double* R, int64_t ldr, double* out, mfx_stream stream);
/*
 * Counts elements with percentDiff(R, X) > threshold using the reference's
 * verify rule (polybench/common/polybenchUtilFuncts.h:27-42: both |v| < 0.01
 * pass, else 100*|R-X|/|R+1e-8| in float).  Blocking: count written to *out.
 */
mfx_status mfx_percent_diff_count(mfx_handle handle, int64_t rows, int64_t cols,
    const float* X, int64_t ldx, const double* R, int64_t ldr, double threshold,
    int64_t* out, mfx_stream stream);

/* ---- input generation and element-wise utilities ------------------------- */

/*
 * p[r*ld + c] = U[-1,1) from SplitMix64 keyed by (seed, matrix_id, r*cols + c):
 * 2 * (h >> 40) * 2^-24 - 1, bit-identical to the host generator.
 */
mfx_status mfx_fill_uniform(mfx_handle handle, int64_t rows, int64_t cols, float* p,
    int64_t ld, uint64_t seed, uint32_t matrix_id, mfx_stream stream);
/* p = sign * exp(sigma * N(0,1) + mu), signs and normals from the same keyed stream. */
mfx_status mfx_fill_lognormal(mfx_handle handle, int64_t rows, int64_t cols, float* p,
    int64_t ld, uint64_t seed, uint32_t matrix_id, float mu, float sigma, mfx_stream stream);
/* hi[i] = bf16_rne(x[i]), lo[i] = bf16_rne(x[i] - hi[i]) as raw 16-bit patterns. */
mfx_status mfx_split_bf16(mfx_handle handle, int64_t count, const float* x, uint16_t* hi,
    uint16_t* lo, mfx_stream stream);
/*
 * Block-scaled OCP e4m3 quantisation of a rows x cols matrix with one
 * power-of-two scale per block x block tile: scale = 2^e with the smallest e
 * such that amax / 2^e <= 448; q = e4m3_rne(x / scale).  q is row-major
 * (rows*cols bytes), exps holds e per tile (row-major over tiles).
 */
mfx_status mfx_quantize_fp8_block(mfx_handle handle, int64_t rows, int64_t cols,
    const float* x, int64_t ldx, int32_t block, uint8_t* q, int32_t* exps, mfx_stream stream);

/* ---- kernel timing (hipEvent pairs around every launch) ------------------ */

/* Starts recording a hipEvent pair around each kernel the handle launches. */
mfx_status mfx_timing_begin(mfx_handle handle);
/*
 * Stops recording, waits for the events, and reports up to `cap` distinct
 * kernel names with launch counts and summed durations in milliseconds.
 * `names` receives pointers to static strings.  *n_out = number of kernels.
 */
mfx_status mfx_timing_end(mfx_handle handle, int32_t cap, const char** names,
    int64_t* launches, double* total_ms, int32_t* n_out);

/* ---- roofline calibrators ---------------------------------------------------- */

/* Number of waves mfx_mfma_burn runs (two per SIMD on every CU). */
int64_t mfx_mfma_burn_waves(mfx_handle handle);
/*
 * Register-only MFMA loop of `prec`'s matrix instruction (FP32: 32x32x2,
 * split / single BF16: 32x32x16, FP8: scaled 32x32x64 e4m3) on two waves per
 * SIMD, 8 independent accumulators per wave.  random_operands = 0: constant
 * power-of-two operands, `iterations` MFMAs per accumulator (the matrix
 * cores' upper bound: their switching power, and so the clock drop, is
 * minimal).  random_operands = 1 (FP32 and BF16 only): per-lane random
 * operands, 4 * iterations MFMAs per accumulator that cancel exactly, plus
 * one constant one (the power draw of a real contraction).  out[w] receives wave w's checksum,
 * which must equal *expected; *flop is the total work, so flop / kernel time
 * is the measured matrix-core ceiling.  Follows micro/arith.cpp:41-68 (a
 * known-result arithmetic loop) on the matrix cores.
 */
mfx_status mfx_mfma_burn(mfx_handle handle, mfx_precision prec, int64_t iterations, int32_t random_operands,
    float* out, int64_t out_len, double* flop, float* expected, mfx_stream stream);
/*
 * dst = src for `bytes` bytes (16-byte aligned buffers, not overlapping),
 * with 128-bit loads and stores: 2 * bytes / kernel time is the measured HBM
 * ceiling.
 * Follows micro/DRAM.cpp:47-62.
 */
mfx_status mfx_hbm_copy(mfx_handle handle, const void* src, void* dst, int64_t bytes, mfx_stream stream);

/* ---- autotuner -------------------------------------------------------------- */

/*
 * Times every compiled tile configuration for (prec, m, n, k) on scratch
 * buffers, records the fastest in the process-wide table and returns its id.
 * Later gemm / grouped calls of the same precision and shape bucket use it.
 */
mfx_status mfx_autotune(mfx_handle handle, mfx_precision prec, int64_t m, int64_t n,
    int64_t k, int32_t* best_config, mfx_stream stream);
/*
 * Tunes the tile policy of a grouped launch on the caller's own problem list:
 * times the launch under every policy (see mfx_grouped_policy_info), records
 * the fastest for launches with the same key (precision, problem count,
 * contraction volume and dependency depth, in power-of-two buckets) and
 * returns its id.  Each policy's launch is planned once and timed over
 * back-to-back launches of the same tables (as a captured graph replays it),
 * in two rounds over the policies, the faster kept.  Runs the launch several
 * times on the given buffers (C is overwritten).  Blocking.
 */
mfx_status mfx_autotune_grouped(mfx_handle handle, mfx_precision prec,
    const mfx_gemm_problem* problems, int32_t count, const mfx_dep* deps, int32_t ndeps,
    int32_t* best_policy, mfx_stream stream);
/*
 * Grouped tile policies: a launch runs one policy, a pair of tile
 * configurations (ids as in mfx_config_info); each problem takes the large
 * one unless that pads its output area by more than `waste_bound` (0: always
 * the large one), else the small one, and on a dependency level whose large
 * tiles would not give every CU one, the small one.  FP32 launches run
 * configuration 3 in place of 0.  *per_level = 1: the launch is split into
 * one pre-pass + persistent kernel per dependency level (kernel boundaries
 * order the levels; no in-kernel waits), else one persistent kernel with
 * in-kernel dependency waits.
 */
int32_t mfx_grouped_policy_count(void);
mfx_status mfx_grouped_policy_info(int32_t policy, int32_t* large_config, int32_t* small_config,
    double* waste_bound, int32_t* per_level);
/* Forces `policy` (or -1 for the tuned / estimated choice) for later grouped launches. */
mfx_status mfx_force_grouped_policy(mfx_handle handle, int32_t policy);
/* Saves / loads the tuning table as JSON. */
mfx_status mfx_autotune_save(const char* path);
mfx_status mfx_autotune_load(const char* path);
/* Number of compiled tile configurations for `prec` and the shape of one. */
int32_t mfx_config_count(mfx_precision prec);
mfx_status mfx_config_info(mfx_precision prec, int32_t config, int32_t* bm, int32_t* bn,
    int32_t* bk, int32_t* waves);
/* Forces `config` (or -1 for the tuned / heuristic choice) for later calls of `prec`. */
mfx_status mfx_force_config(mfx_handle handle, mfx_precision prec, int32_t config);
/* FP32 split-K override for later single-product calls: -1 automatic (the
 * default: k slices when the tile grid leaves the CUs short of work), 1 never,
 * >= 2 that many k slices (fewer when K or the 32 MiB partial-tile bound does
 * not allow them; configurations with tiles above 128x128 never split).
 * Split results are summed in slice order, so they are reproducible, but not
 * bit-identical to the unsplit kernel's.  No counterpart in the reference. */
mfx_status mfx_force_splits(mfx_handle handle, int32_t splits);

#ifdef __cplusplus
}
#endif

#endif /* MFX_H */
