// The C ABI of include/mfx.h: argument validation, handle and workspace
// management, kernel timing, and dispatch to the kernel families.
#include <algorithm>
#include <cstdio>
#include <map>

#include "tile_bf16.h"

namespace mfx {

// Kernel-family entry points defined in the other translation units.
int f32_config_count();
bool f32_config_info(int c, int* bm, int* bn, int* bk, int* waves);
int bf_config_count();
bool bf_config_info(int c, bool split, int* bm, int* bn, int* bk, int* waves);
int fp8_config_count();
bool fp8_config_info(int c, int* bm, int* bn, int* bk, int* waves);
mfx_status launch_gemm_f64(Handle* h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
    int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s);
mfx_status launch_convert(Handle* h, int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst,
    int64_t ldd, hipStream_t s);
mfx_status launch_rel_frobenius(Handle* h, int uplo, int64_t rows, int64_t cols, const float* X, int64_t ldx,
    const double* R, int64_t ldr, double* out_host, hipStream_t s);
mfx_status launch_percent_diff(Handle* h, int64_t rows, int64_t cols, const float* X, int64_t ldx, const double* R,
    int64_t ldr, double threshold, int64_t* out_host, hipStream_t s);
mfx_status launch_fill_uniform(Handle* h, int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed,
    uint32_t mid, hipStream_t s);
mfx_status launch_fill_lognormal(Handle* h, int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed,
    uint32_t mid, float mu, float sigma, hipStream_t s);
mfx_status launch_split_elementwise(Handle* h, int64_t n, const float* x, uint16_t* hi, uint16_t* lo, hipStream_t s);
mfx_status launch_quantize_fp8(Handle* h, int64_t rows, int64_t cols, const float* x, int64_t ldx, int block,
    uint8_t* q, int32_t* exps, hipStream_t s);

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

mfx_status hip_fail(hipError_t e, const char* what) {
  set_last_error(std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
  if(e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return MFX_ERROR_ALLOC;
  if(e == hipErrorNoDevice || e == hipErrorInvalidDevice) return MFX_ERROR_NO_DEVICE;
  return MFX_ERROR_HIP;
}

mfx_status invalid(const std::string& msg) {
  set_last_error(msg);
  return MFX_ERROR_INVALID_VALUE;
}

mfx_status unsupported(const std::string& msg) {
  set_last_error(msg);
  return MFX_ERROR_NOT_SUPPORTED;
}

hipEvent_t take_event(Handle* h) {
  if(!h->free_events.empty()) {
    hipEvent_t e = h->free_events.back();
    h->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if(hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

mfx_status workspace(Handle* h, size_t bytes, void** out) {
  if(bytes <= h->workspace_bytes && h->workspace) {
    *out = h->workspace;
    return MFX_SUCCESS;
  }
  if(h->capturing) return invalid("workspace must be reserved before graph capture");
  if(h->workspace) {
    MFX_HIP(hipDeviceSynchronize());
    MFX_HIP(hipFree(h->workspace));
    h->workspace = nullptr;
    h->workspace_bytes = 0;
  }
  const size_t want = (size_t)round_up((int64_t)bytes, 1 << 21);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if(e != hipSuccess) return hip_fail(e, "hipMalloc(workspace)");
  h->workspace = p;
  h->workspace_bytes = want;
  *out = p;
  return MFX_SUCCESS;
}

int config_count(mfx_precision p) {
  switch(p) {
    case MFX_PREC_FP32: return f32_config_count();
    case MFX_PREC_SPLIT_BF16:
    case MFX_PREC_BF16: return bf_config_count();
    case MFX_PREC_FP8_BLOCK:
    case MFX_PREC_FP8_BLOCK2: return fp8_config_count();
  }
  return 0;
}

bool config_info(mfx_precision p, int c, int* bm, int* bn, int* bk, int* waves) {
  switch(p) {
    case MFX_PREC_FP32: return f32_config_info(c, bm, bn, bk, waves);
    case MFX_PREC_SPLIT_BF16: return bf_config_info(c, true, bm, bn, bk, waves);
    case MFX_PREC_BF16: return bf_config_info(c, false, bm, bn, bk, waves);
    case MFX_PREC_FP8_BLOCK:
    case MFX_PREC_FP8_BLOCK2: return fp8_config_info(c, bm, bn, bk, waves);
  }
  return false;
}

// Tuned choices: (precision, shape bucket) -> config.  Buckets are log2 of
// m, n, k rounded up, so shapes within a factor of two share a decision.
namespace {
std::mutex g_tune_mutex;
std::map<std::string, int> g_tuned;

int log2_bucket(int64_t x) {
  int b = 0;
  while((int64_t(1) << b) < x && b < 40) ++b;
  return b;
}
}  // namespace

std::string tune_key(mfx_precision p, int64_t m, int64_t n, int64_t k) {
  char buf[64];
  snprintf(buf, sizeof buf, "%d:%d:%d:%d", (int)p, log2_bucket(m), log2_bucket(n), log2_bucket(k));
  return buf;
}

void record_tuned(const std::string& key, int config) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tuned[key] = config;
}

std::map<std::string, int> tuned_snapshot() {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  return g_tuned;
}

void replace_tuned(const std::map<std::string, int>& t) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tuned = t;
}

// Per-flop rate of each tile configuration relative to the best one, on
// large problems (the best configuration's time over this one's, measured on
// MI355X at M = N = K = 8192).
double config_rate(mfx_precision p, int c) {
  static const double f32[] = {0.886, 0.868, 0.817, 0.859, 1.0, 0.97, 0.913, 0.81};
  static const double split[] = {0.894, 0.846, 0.645, 0.70, 1.0, 0.80, 0.854};
  static const double bf16[] = {0.844, 0.730, 0.508, 0.662, 1.0, 0.724, 0.814};
  static const double fp8[] = {0.866, 0.69, 1.0};
  const double* t = p == MFX_PREC_FP32 ? f32 : p == MFX_PREC_SPLIT_BF16 ? split : p == MFX_PREC_BF16 ? bf16 : fp8;
  const int n = p == MFX_PREC_FP32 ? 8 : p == MFX_PREC_SPLIT_BF16 || p == MFX_PREC_BF16 ? 7 : 3;
  return c >= 0 && c < n ? t[c] : 0.5;
}

// Heuristic used when nothing is tuned: the configuration with the least
// estimated time, counted in rounds of tiles over the CUs (a round costs the
// tile's area over its rate), so a last round that leaves most CUs idle is
// paid for.  `tiles(bm, bn)` gives the tile count of the schedule (the full
// grid, or a triangle).  Ties go to the larger tile, then more waves.
// Exception for the power-limited BF16 / FP8 MFMAs: when the largest tile
// fills between half and all of the CUs in one round, it is taken (the clock
// rises as CUs idle, which the round count does not see; measured on MI355X,
// e.g. a 4096 split-BF16 syrk: 136 large tiles beat 528 small ones by 10 %).
template <class Tiles>
int least_time_config(Handle* h, mfx_precision p, bool square_only, Tiles&& tiles) {
  const int cus = h->cu_count > 0 ? h->cu_count : 256;
  if(p != MFX_PREC_FP32) {
    int big = -1, big_w = 0;
    int64_t big_area = 0;
    for(int c = 0; c < config_count(p); ++c) {
      int bm, bn, bk, w;
      if(!config_info(p, c, &bm, &bn, &bk, &w) || (square_only && bm != bn)) continue;
      const int64_t area = (int64_t)bm * bn;
      if(area > big_area || (area == big_area && w > big_w)) {
        big = c;
        big_area = area;
        big_w = w;
      }
    }
    if(big >= 0) {
      int bm, bn, bk, w;
      config_info(p, big, &bm, &bn, &bk, &w);
      const int64_t t = tiles(bm, bn);
      if(2 * t >= cus && t <= cus) return big;
    }
  }
  int best = -1;
  double best_t = 0.0;
  int64_t best_area = 0;
  int best_waves = 0;
  for(int c = 0; c < config_count(p); ++c) {
    int bm, bn, bk, w;
    if(!config_info(p, c, &bm, &bn, &bk, &w) || (square_only && bm != bn)) continue;
    const int64_t area = (int64_t)bm * bn;
    const double t = (double)ceil_div(tiles(bm, bn), cus) * (double)area / config_rate(p, c);
    const bool better = best < 0 || t < best_t * (1.0 - 1e-9) ||
                        (t <= best_t * (1.0 + 1e-9) && (area > best_area || (area == best_area && w > best_waves)));
    if(better) {
      best = c;
      best_t = t;
      best_area = area;
      best_waves = w;
    }
  }
  return best < 0 ? 0 : best;
}

// FP32 configuration with sixteen waves on a 64x64 tile (four k-groups).
constexpr int kF32KGroupConfig = 7;

int heuristic_config(Handle* h, mfx_precision p, int64_t m, int64_t n, int64_t k) {
  (void)k;
  // FP32 grids of at most one 64x64 tile per CU: the k-grouped tile (one
  // sixteen-wave workgroup per CU hides load latency that four-wave tiles
  // leave exposed at this size; split-K then fills the CUs if tiles are
  // few).  Measured on MI355X: 1024^3 in 27-30 us against 32-36 us.
  const int cus = h->cu_count > 0 ? h->cu_count : 256;
  if(p == MFX_PREC_FP32 && ceil_div(m, 64) * ceil_div(n, 64) <= cus) return kF32KGroupConfig;
  return least_time_config(h, p, false, [&](int bm, int bn) { return ceil_div(m, bm) * ceil_div(n, bn); });
}

int choose_config(Handle* h, mfx_precision p, int64_t m, int64_t n, int64_t k) {
  if(h->forced_config[p] >= 0) return h->forced_config[p];
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tuned.find(tune_key(p, m, n, k));
    if(it != g_tuned.end()) return it->second;
  }
  return heuristic_config(h, p, m, n, k);
}

// Tuned configuration (or grouped tile policy) recorded under `key`, -1 if none.
int tuned_lookup(const std::string& key) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  auto it = g_tuned.find(key);
  return it == g_tuned.end() ? -1 : it->second;
}

// The square configuration a triangle's sparse last round is re-tiled with:
// half or a quarter of the big tile's side.
static int tri_small_config(mfx_precision p, int side) {
  if(p == MFX_PREC_FP32) return side == 128 ? 6 : side == 64 ? 7 : -1;
  if(p == MFX_PREC_SPLIT_BF16 || p == MFX_PREC_BF16) return side == 128 ? 6 : side == 64 ? 2 : -1;
  return -1;
}

TriTail tri_tail(mfx_precision p, int config, int64_t T, int cus) {
  TriTail t;
  int bm, bn, bk, w;
  if(!config_info(p, config, &bm, &bn, &bk, &w) || bm != bn || cus <= 0) return t;
  const int64_t tiles = T * (T + 1) / 2, rounds = tiles / cus, tail = tiles - rounds * cus;
  t.head = (int)tiles;
  if(rounds == 0 || tail == 0 || 2 * tail > cus) return t;
  // one more round of big tiles, against ceil(tail r^2 / cus) rounds of small ones
  double best = (double)bm * bm / config_rate(p, config);
  for(int r : {2, 4}) {
    const int s = tri_small_config(p, bm / r);
    if(s < 0) continue;
    const double cost = (double)ceil_div(tail * r * r, cus) * (double)(bm / r) * (bm / r) / config_rate(p, s);
    if(cost < best) {
      best = cost;
      t.small = s;
      t.ratio = r;
    }
  }
  if(t.small >= 0) t.head = (int)(rounds * cus);
  return t;
}

// Least estimated time over the square configurations for an n x n
// triangle: rounds of T(T+1)/2 tiles over the CUs, each costing the tile's
// area over its rate, with a sparse last round re-tiled as tri_tail does.
// BF16 / FP8 take their largest tile when it fills between half and all of
// the CUs in one round (the power-limited clock rises as CUs idle; see
// least_time_config).
static int least_time_triangle(Handle* h, mfx_precision p, int64_t n) {
  const int cus = h->cu_count > 0 ? h->cu_count : 256;
  auto tri = [&](int b) {
    const int64_t t = ceil_div(n, b);
    return t * (t + 1) / 2;
  };
  int best = -1, best_w = 0;
  double best_t = 0.0;
  int64_t best_area = 0;
  if(p != MFX_PREC_FP32) {
    for(int c = 0; c < config_count(p); ++c) {
      int bm, bn, bk, w;
      if(!config_info(p, c, &bm, &bn, &bk, &w) || bm != bn) continue;
      if((int64_t)bm * bn > best_area || ((int64_t)bm * bn == best_area && w > best_w)) {
        best = c;
        best_area = (int64_t)bm * bn;
        best_w = w;
      }
    }
    if(best >= 0) {
      int bm, bn, bk, w;
      config_info(p, best, &bm, &bn, &bk, &w);
      const int64_t t = tri(bm);
      if(2 * t >= cus && t <= cus) return best;
    }
    best = -1;
    best_area = 0;
    best_w = 0;
  }
  for(int c = 0; c < config_count(p); ++c) {
    int bm, bn, bk, w;
    if(!config_info(p, c, &bm, &bn, &bk, &w) || bm != bn) continue;
    const int64_t T = ceil_div(n, bm), tiles = tri(bm), area = (int64_t)bm * bn;
    const TriTail tt = tri_tail(p, c, T, cus);
    double t;
    if(tt.small >= 0) {
      const int sb = bm / tt.ratio;
      t = (double)(tiles / cus) * area / config_rate(p, c) +
          (double)ceil_div((tiles - tt.head) * tt.ratio * tt.ratio, cus) * sb * sb / config_rate(p, tt.small);
    } else {
      t = (double)ceil_div(tiles, cus) * area / config_rate(p, c);
    }
    const bool better = best < 0 || t < best_t * (1.0 - 1e-9) ||
                        (t <= best_t * (1.0 + 1e-9) && (area > best_area || (area == best_area && w > best_w)));
    if(better) {
      best = c;
      best_t = t;
      best_area = area;
      best_w = w;
    }
  }
  return best < 0 ? 0 : best;
}

// Square-tile choice for the triangular schedules: a forced or tuned square
// configuration, else least_time_triangle.
int choose_square_config(Handle* h, mfx_precision p, int64_t n, int64_t k) {
  int c = -1;
  if(h->forced_config[p] >= 0) {
    c = h->forced_config[p];
  } else {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tuned.find(tune_key(p, n, n, k));
    if(it != g_tuned.end()) c = it->second;
  }
  int bm, bn, bk, w;
  if(c >= 0) return config_info(p, c, &bm, &bn, &bk, &w) && bm == bn ? c : 0;
  return least_time_triangle(h, p, n);
}

mfx_status dispatch_dense(Handle* h, mfx_precision prec, DenseOp& op, int config, hipStream_t s) {
  switch(prec) {
    case MFX_PREC_FP32: return launch_dense_f32(h, op, config, s);
    case MFX_PREC_SPLIT_BF16: return launch_dense_bf16x(h, op, true, config, s);
    case MFX_PREC_BF16: return launch_dense_bf16x(h, op, false, config, s);
    case MFX_PREC_FP8_BLOCK: return launch_dense_fp8(h, op, false, config, s);
    case MFX_PREC_FP8_BLOCK2: return launch_dense_fp8(h, op, true, config, s);
  }
  return invalid("unknown precision");
}

bool valid_prec(int p) { return p >= MFX_PREC_FP32 && p <= MFX_PREC_FP8_BLOCK2; }

// Shared validation of one row-major operand view.
mfx_status check_matrix(const char* name, const void* p, int64_t rows, int64_t cols, int64_t ld) {
  if(rows < 0 || cols < 0) return invalid(std::string(name) + ": negative dimension");
  if(rows > 0 && cols > 0 && p == nullptr) return invalid(std::string(name) + ": null pointer");
  if(ld < std::max<int64_t>(1, cols)) return invalid(std::string(name) + ": leading dimension smaller than row length");
  if(rows > (int64_t)1 << 30 || cols > (int64_t)1 << 30) return invalid(std::string(name) + ": dimension too large");
  return MFX_SUCCESS;
}

// K limit of the packed precisions (32-bit panel offsets in the BF16 tile).
mfx_status check_k(mfx_precision prec, int64_t k) {
  if((prec == MFX_PREC_SPLIT_BF16 || prec == MFX_PREC_BF16) && !bf16_k_supported(k, prec == MFX_PREC_SPLIT_BF16 ? 2 : 1))
    return unsupported("K = " + std::to_string(k) + " is beyond the packed BF16 path's limit (2^25 split, 2^26 single)");
  return MFX_SUCCESS;
}

// Argument checks of mfx_gemm; pointers are checked only when check_ptrs.
mfx_status check_gemm_args(mfx_precision prec, mfx_op transA, mfx_op transB, int64_t m, int64_t n, int64_t k,
    const void* A, int64_t lda, const void* B, int64_t ldb, const void* C, int64_t ldc) {
  if(!valid_prec(prec)) return invalid("unknown precision");
  if((transA != MFX_OP_N && transA != MFX_OP_T) || (transB != MFX_OP_N && transB != MFX_OP_T))
    return invalid("transA / transB must be MFX_OP_N or MFX_OP_T");
  if(m < 0 || n < 0 || k < 0) return invalid("negative dimension");
  MFX_CHECK(transA == MFX_OP_N ? check_matrix("A", A, m, k, lda) : check_matrix("A", A, k, m, lda));
  MFX_CHECK(transB == MFX_OP_N ? check_matrix("B", B, k, n, ldb) : check_matrix("B", B, n, k, ldb));
  MFX_CHECK(check_matrix("C", C, m, n, ldc));
  if(m == 0 || n == 0) return MFX_SUCCESS;
  return check_k(prec, k);
}

// C (n x n) = alpha * A·Aᵀ (syr2k: A·Bᵀ + B·Aᵀ) + beta * C on the
// triangle schedule.  A is n x k row-major, or with `trans` k x n (so the
// product is Aᵀ·A, the covariance form).
mfx_status sym_common(Handle* h, mfx_precision prec, mfx_uplo uplo, int mirror, int64_t n, int64_t k, float alpha,
    const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int nprod, bool trans,
    hipStream_t s) {
  if(!valid_prec(prec)) return invalid("unknown precision");
  if(uplo != MFX_LOWER && uplo != MFX_UPPER && uplo != MFX_FULL) return invalid("uplo must be LOWER, UPPER or FULL");
  if(n < 0 || k < 0) return invalid("negative dimension");
  MFX_CHECK(trans ? check_matrix("A", A, k, n, lda) : check_matrix("A", A, n, k, lda));
  if(nprod == 2) MFX_CHECK(trans ? check_matrix("B", B, k, n, ldb) : check_matrix("B", B, n, k, ldb));
  MFX_CHECK(check_matrix("C", C, n, n, ldc));
  if(n == 0) return MFX_SUCCESS;
  MFX_CHECK(check_k(prec, k));
  DenseOp op{};
  op.M = op.N = (int)n;
  op.K = (int)k;
  op.A = A;
  op.lda = lda;
  op.a_kmajor = !trans;
  op.B = nprod == 2 ? B : A;
  op.ldb = nprod == 2 ? ldb : lda;
  op.b_kmajor = !trans;
  op.nprod = nprod;
  const int mode = uplo == MFX_LOWER ? TILES_LOWER : uplo == MFX_UPPER ? TILES_UPPER : TILES_FULL;
  op.tiles = mode;
  op.ep = Epilogue{C, ldc, alpha, beta, (int)n, (int)n, mode, mode == TILES_FULL ? 0 : (mirror ? 1 : 0)};
  const int cfg = mode == TILES_FULL ? choose_config(h, prec, n, n, k * nprod) : choose_square_config(h, prec, n, k * nprod);
  return dispatch_dense(h, prec, op, cfg, s);
}

}  // namespace mfx

using namespace mfx;

#define HANDLE_OR_FAIL(hh)                                          \
  if(!(hh)) return invalid("null handle");                          \
  Handle* h = &(hh)->h;                                             \
  if(hipSetDevice(h->device) != hipSuccess) return hip_fail(hipErrorInvalidDevice, "hipSetDevice");

extern "C" {

mfx_status mfx_create(mfx_handle* handle, int device) {
  if(!handle) return invalid("null handle pointer");
  *handle = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if(e != hipSuccess || n == 0) {
    set_last_error("no HIP device available");
    return MFX_ERROR_NO_DEVICE;
  }
  if(device < 0 || device >= n) return invalid("device index out of range");
  MFX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MFX_HIP(hipGetDeviceProperties(&prop, device));
  if(std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) {
    set_last_error(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950 only");
    return MFX_ERROR_NO_DEVICE;
  }
  MFX_CHECK(noalloc_init(device));
  auto* hs = new(std::nothrow) mfx_handle_s;
  if(!hs) return MFX_ERROR_ALLOC;
  hs->h.device = device;
  hs->h.cu_count = prop.multiProcessorCount;
  if(const char* sl = std::getenv("MFX_SPIN_LIMIT")) hs->h.spin_limit = std::max(0, std::atoi(sl));
  hs->h.device_name = prop.name;
  // The marketing name can be empty when the driver's id table is missing.
  if(hs->h.device_name.empty()) hs->h.device_name = std::string("AMD GPU ") + prop.gcnArchName;
  *handle = hs;
  return MFX_SUCCESS;
}

mfx_status mfx_destroy(mfx_handle handle) {
  if(!handle) return MFX_SUCCESS;
  Handle* h = &handle->h;
  // Teardown continues past errors; the first one is reported.
  hipError_t first = hipSetDevice(h->device);
  auto keep = [&first](hipError_t e) {
    if(first == hipSuccess) first = e;
  };
  keep(hipDeviceSynchronize());
  for(auto& r : h->records) {
    keep(hipEventDestroy(r.start));
    keep(hipEventDestroy(r.stop));
  }
  for(auto e : h->free_events) keep(hipEventDestroy(e));
  if(h->workspace) keep(hipFree(h->workspace));
  if(h->staging) keep(hipHostFree(h->staging));
  if(h->staging_ready) keep(hipEventDestroy(h->staging_ready));
  if(h->error_flag) keep(hipFree(h->error_flag));
  delete handle;
  return first == hipSuccess ? MFX_SUCCESS : hip_fail(first, "mfx_destroy");
}

mfx_status mfx_reserve_workspace(mfx_handle handle, size_t bytes) {
  HANDLE_OR_FAIL(handle);
  void* p;
  return workspace(h, bytes, &p);
}

mfx_status mfx_gemm_check(mfx_precision prec, mfx_op transA, mfx_op transB, int64_t m, int64_t n, int64_t k,
    int64_t lda, int64_t ldb, int64_t ldc) {
  // any non-null address stands in for the operands: only shapes are checked
  static const float stand_in = 0.0f;
  return check_gemm_args(prec, transA, transB, m, n, k, &stand_in, lda, &stand_in, ldb, &stand_in, ldc);
}

size_t mfx_gemm_workspace_size(mfx_precision prec, int64_t m, int64_t n, int64_t k) {
  if(!valid_prec(prec) || m <= 0 || n <= 0 || k <= 0) return 0;
  DenseOp op{};
  op.M = (int)m;
  op.N = (int)n;
  op.K = (int)k;
  op.nprod = 1;
  size_t best = 0;
  // FP32 needs workspace only for split-K partial tiles, whose count depends
  // on the CUs of the current device.
  int cus = 0;
  if(prec == MFX_PREC_FP32) {
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  // The largest need over all configurations, so any tuned choice fits.
  for(int c = 0; c < config_count(prec); ++c) {
    size_t b = 0;
    if(prec == MFX_PREC_FP32) {
      int bm, bn, bk, w;
      config_info(prec, c, &bm, &bn, &bk, &w);
      const int64_t tiles = ceil_div(m, bm) * ceil_div(n, bn);
      const int splits = f32_splits(c, tiles, k, cus);
      if(splits > 1) b = f32_splitk_workspace(c, tiles, splits);
    } else if(prec == MFX_PREC_SPLIT_BF16 || prec == MFX_PREC_BF16)
      b = dense_bf16x_workspace(op, prec == MFX_PREC_SPLIT_BF16, c);
    else if(prec == MFX_PREC_FP8_BLOCK || prec == MFX_PREC_FP8_BLOCK2)
      b = dense_fp8_workspace(op, prec == MFX_PREC_FP8_BLOCK2, c);
    best = std::max(best, b);
  }
  return best;
}

const char* mfx_status_string(mfx_status status) {
  switch(status) {
    case MFX_SUCCESS: return "MFX_SUCCESS";
    case MFX_ERROR_INVALID_VALUE: return "MFX_ERROR_INVALID_VALUE";
    case MFX_ERROR_NOT_SUPPORTED: return "MFX_ERROR_NOT_SUPPORTED";
    case MFX_ERROR_HIP: return "MFX_ERROR_HIP";
    case MFX_ERROR_ALLOC: return "MFX_ERROR_ALLOC";
    case MFX_ERROR_NO_DEVICE: return "MFX_ERROR_NO_DEVICE";
    case MFX_ERROR_TIMEOUT: return "MFX_ERROR_TIMEOUT";
    case MFX_ERROR_IO: return "MFX_ERROR_IO";
  }
  return "MFX_ERROR_UNKNOWN";
}

const char* mfx_last_error(void) { return g_last_error.c_str(); }

mfx_status mfx_device_name(mfx_handle handle, char* buf, size_t len) {
  HANDLE_OR_FAIL(handle);
  if(!buf || len == 0) return invalid("null buffer");
  snprintf(buf, len, "%s", h->device_name.c_str());
  return MFX_SUCCESS;
}

mfx_status mfx_gemm(mfx_handle handle, mfx_precision prec, mfx_op transA, mfx_op transB, int64_t m, int64_t n,
    int64_t k, float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc,
    mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  MFX_CHECK(check_gemm_args(prec, transA, transB, m, n, k, A, lda, B, ldb, C, ldc));
  if(m == 0 || n == 0) return MFX_SUCCESS;
  DenseOp op{};
  op.M = (int)m;
  op.N = (int)n;
  op.K = (int)k;
  op.A = A;
  op.lda = lda;
  op.a_kmajor = transA == MFX_OP_N;
  op.B = B;
  op.ldb = ldb;
  op.b_kmajor = transB == MFX_OP_T;
  op.nprod = 1;
  op.tiles = TILES_FULL;
  op.ep = Epilogue{C, ldc, alpha, beta, (int)m, (int)n, TILES_FULL, 0};
  if(k == 0) {
    // C = beta * C (alpha * empty sum): the contraction loop runs zero times.
    op.K = 0;
  }
  const int cfg = choose_config(h, prec, m, n, k);
  return dispatch_dense(h, prec, op, cfg, static_cast<hipStream_t>(stream));
}


mfx_status mfx_syrk(mfx_handle handle, mfx_precision prec, mfx_uplo uplo, int mirror, int64_t n, int64_t k, float alpha,
    const float* A, int64_t lda, float beta, float* C, int64_t ldc, mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  return sym_common(h, prec, uplo, mirror, n, k, alpha, A, lda, nullptr, 0, beta, C, ldc, 1, false,
      static_cast<hipStream_t>(stream));
}

mfx_status mfx_syr2k(mfx_handle handle, mfx_precision prec, mfx_uplo uplo, int mirror, int64_t n, int64_t k,
    float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc,
    mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  return sym_common(h, prec, uplo, mirror, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 2, false,
      static_cast<hipStream_t>(stream));
}

mfx_status mfx_gemm_f64(mfx_handle handle, mfx_op transA, mfx_op transB, int64_t m, int64_t n, int64_t k, double alpha,
    const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
    mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  if((transA != MFX_OP_N && transA != MFX_OP_T) || (transB != MFX_OP_N && transB != MFX_OP_T))
    return invalid("transA / transB must be MFX_OP_N or MFX_OP_T");
  if(m < 0 || n < 0 || k < 0) return invalid("negative dimension");
  MFX_CHECK(transA == MFX_OP_N ? check_matrix("A", A, m, k, lda) : check_matrix("A", A, k, m, lda));
  MFX_CHECK(transB == MFX_OP_N ? check_matrix("B", B, k, n, ldb) : check_matrix("B", B, n, k, ldb));
  MFX_CHECK(check_matrix("C", C, m, n, ldc));
  return launch_gemm_f64(h, transA == MFX_OP_T, transB == MFX_OP_T, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
      static_cast<hipStream_t>(stream));
}

mfx_status mfx_convert_f32_f64(mfx_handle handle, int64_t rows, int64_t cols, const float* src, int64_t lds,
    double* dst, int64_t ldd, mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  MFX_CHECK(check_matrix("src", src, rows, cols, lds));
  MFX_CHECK(check_matrix("dst", dst, rows, cols, ldd));
  return launch_convert(h, rows, cols, src, lds, dst, ldd, static_cast<hipStream_t>(stream));
}

mfx_status mfx_rel_frobenius(mfx_handle handle, mfx_uplo uplo, int64_t rows, int64_t cols, const float* X, int64_t ldx,
    const double* R, int64_t ldr, double* out, mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  if(!out) return invalid("null output pointer");
  if(uplo != MFX_LOWER && uplo != MFX_UPPER && uplo != MFX_FULL) return invalid("bad uplo");
  MFX_CHECK(check_matrix("X", X, rows, cols, ldx));
  MFX_CHECK(check_matrix("R", R, rows, cols, ldr));
  return launch_rel_frobenius(h, uplo, rows, cols, X, ldx, R, ldr, out, static_cast<hipStream_t>(stream));
}

mfx_status mfx_percent_diff_count(mfx_handle handle, int64_t rows, int64_t cols, const float* X, int64_t ldx,
    const double* R, int64_t ldr, double threshold, int64_t* out, mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  if(!out) return invalid("null output pointer");
  MFX_CHECK(check_matrix("X", X, rows, cols, ldx));
  MFX_CHECK(check_matrix("R", R, rows, cols, ldr));
  return launch_percent_diff(h, rows, cols, X, ldx, R, ldr, threshold, out, static_cast<hipStream_t>(stream));
}

mfx_status mfx_fill_uniform(mfx_handle handle, int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed,
    uint32_t matrix_id, mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  MFX_CHECK(check_matrix("p", p, rows, cols, ld));
  return launch_fill_uniform(h, rows, cols, p, ld, seed, matrix_id, static_cast<hipStream_t>(stream));
}

mfx_status mfx_fill_lognormal(mfx_handle handle, int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed,
    uint32_t matrix_id, float mu, float sigma, mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  MFX_CHECK(check_matrix("p", p, rows, cols, ld));
  if(!(sigma >= 0.0f)) return invalid("sigma must be >= 0");
  return launch_fill_lognormal(h, rows, cols, p, ld, seed, matrix_id, mu, sigma, static_cast<hipStream_t>(stream));
}

mfx_status mfx_split_bf16(mfx_handle handle, int64_t count, const float* x, uint16_t* hi, uint16_t* lo,
    mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  if(count < 0) return invalid("negative count");
  if(count > 0 && (!x || !hi || !lo)) return invalid("null pointer");
  return launch_split_elementwise(h, count, x, hi, lo, static_cast<hipStream_t>(stream));
}

mfx_status mfx_quantize_fp8_block(mfx_handle handle, int64_t rows, int64_t cols, const float* x, int64_t ldx,
    int32_t block, uint8_t* q, int32_t* exps, mfx_stream stream) {
  HANDLE_OR_FAIL(handle);
  MFX_CHECK(check_matrix("x", x, rows, cols, ldx));
  if(block != 128 && block != 64 && block != 32) return invalid("block must be 32, 64 or 128");
  if(rows > 0 && cols > 0 && (!q || !exps)) return invalid("null output pointer");
  return launch_quantize_fp8(h, rows, cols, x, ldx, block, q, exps, static_cast<hipStream_t>(stream));
}

mfx_status mfx_timing_begin(mfx_handle handle) {
  HANDLE_OR_FAIL(handle);
  for(auto& r : h->records) {
    h->free_events.push_back(r.start);
    h->free_events.push_back(r.stop);
  }
  h->records.clear();
  h->timing = true;
  return MFX_SUCCESS;
}

mfx_status mfx_timing_end(mfx_handle handle, int32_t cap, const char** names, int64_t* launches, double* total_ms,
    int32_t* n_out) {
  HANDLE_OR_FAIL(handle);
  h->timing = false;
  std::vector<const char*> order;
  std::map<std::string, std::pair<int64_t, double>> agg;
  for(auto& r : h->records) {
    MFX_HIP(hipEventSynchronize(r.stop));
    float ms = 0.0f;
    MFX_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    auto it = agg.find(r.name);
    if(it == agg.end()) {
      order.push_back(r.name);
      agg[r.name] = {1, (double)ms};
    } else {
      it->second.first += 1;
      it->second.second += ms;
    }
  }
  for(auto& r : h->records) {
    h->free_events.push_back(r.start);
    h->free_events.push_back(r.stop);
  }
  h->records.clear();
  const int n = (int)order.size();
  if(n_out) *n_out = n;
  for(int i = 0; i < n && i < cap; ++i) {
    auto& v = agg[order[i]];
    if(names) names[i] = order[i];
    if(launches) launches[i] = v.first;
    if(total_ms) total_ms[i] = v.second;
  }
  return MFX_SUCCESS;
}

int32_t mfx_config_count(mfx_precision prec) { return valid_prec(prec) ? config_count(prec) : 0; }

mfx_status mfx_config_info(mfx_precision prec, int32_t config, int32_t* bm, int32_t* bn, int32_t* bk, int32_t* waves) {
  if(!valid_prec(prec)) return invalid("unknown precision");
  if(!bm || !bn || !bk || !waves) return invalid("null output pointer");
  int a, b, c, d;
  if(!config_info(prec, config, &a, &b, &c, &d)) return invalid("configuration out of range");
  *bm = a;
  *bn = b;
  *bk = c;
  *waves = d;
  return MFX_SUCCESS;
}

mfx_status mfx_force_config(mfx_handle handle, mfx_precision prec, int32_t config) {
  HANDLE_OR_FAIL(handle);
  if(!valid_prec(prec)) return invalid("unknown precision");
  if(config < -1 || config >= config_count(prec)) return invalid("configuration out of range");
  h->forced_config[prec] = config;
  return MFX_SUCCESS;
}

mfx_status mfx_force_splits(mfx_handle handle, int32_t splits) {
  HANDLE_OR_FAIL(handle);
  if(splits < -1 || splits == 0 || splits > 64) return invalid("splits must be -1, or 1 to 64");
  h->forced_splits = splits;
  return MFX_SUCCESS;
}

}  // extern "C"
