// Internal declarations shared by the libmfx translation units: the handle,
// launch bookkeeping, number formats (BF16 / OCP e4m3 / keyed RNG) and the
// tile-scheduling helpers used by every kernel family.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/mfx.h"

namespace mfx {

// ---------------------------------------------------------------------------
// Error plumbing
// ---------------------------------------------------------------------------
void set_last_error(const std::string& msg);
mfx_status hip_fail(hipError_t e, const char* what);

#define MFX_HIP(expr)                                                   \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if(_e != hipSuccess) return ::mfx::hip_fail(_e, #expr);             \
  } while(0)

#define MFX_CHECK(expr)                                                 \
  do {                                                                  \
    mfx_status _s = (expr);                                             \
    if(_s != MFX_SUCCESS) return _s;                                    \
  } while(0)

mfx_status invalid(const std::string& msg);
mfx_status unsupported(const std::string& msg);

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct TimingRecord {
  const char* name;
  hipEvent_t start;
  hipEvent_t stop;
};

struct Handle {
  int device = 0;
  int cu_count = 0;
  std::string device_name;
  void* workspace = nullptr;
  size_t workspace_bytes = 0;
  bool timing = false;
  std::vector<TimingRecord> records;
  std::vector<hipEvent_t> free_events;
  int forced_config[5] = {-1, -1, -1, -1, -1};
  bool capturing = false;  // set while a grouped launch is being captured
  // Grouped launches: spin limit of the in-kernel dependency waits (env
  // MFX_SPIN_LIMIT overrides), pinned staging for the work tables and the
  // event that marks their upload, and the error flag of the last launch.
  int spin_limit = 1 << 22;
  int grouped_policy = -1;  // forced grouped tile policy (-1: tuned / estimated)
  void* staging = nullptr;
  size_t staging_bytes = 0;
  hipEvent_t staging_ready = nullptr;
  int* error_flag = nullptr;
  int forced_splits = -1;  // FP32 split-K: -1 automatic, 1 never, >= 2 that many slices
};

// Looks up the device's fixed no-allocation scratch array (mfx_create calls
// it with the device current).
mfx_status noalloc_init(int device);

// Returns a workspace pointer of at least `bytes` (grows the arena; growing
// synchronises the device, so it must not happen under graph capture).
mfx_status workspace(Handle* h, size_t bytes, void** out);

// Records an event pair around `launch()` when timing is on, and converts a
// launch error into MFX_ERROR_HIP.
template <class F>
mfx_status timed_launch(Handle* h, const char* name, hipStream_t s, F&& launch);

hipEvent_t take_event(Handle* h);

template <class F>
mfx_status timed_launch(Handle* h, const char* name, hipStream_t s, F&& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  if(h->timing) {
    a = take_event(h);
    b = take_event(h);
    if(!a || !b) return hip_fail(hipErrorOutOfMemory, "hipEventCreate");
    MFX_HIP(hipEventRecord(a, s));
  }
  launch();
  hipError_t e = hipGetLastError();
  if(e != hipSuccess) return hip_fail(e, name);
  if(h->timing) {
    MFX_HIP(hipEventRecord(b, s));
    h->records.push_back({name, a, b});
  }
  return MFX_SUCCESS;
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
inline int64_t ceil_div(int64_t x, int64_t m) { return (x + m - 1) / m; }

// ---------------------------------------------------------------------------
// Number formats (host + device, bit-exact on both sides)
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t f32_bits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}
__host__ __device__ inline float bits_f32(uint32_t u) {
  float x;
  memcpy(&x, &u, 4);
  return x;
}

// Round-to-nearest-even FP32 -> BF16 (NaN stays NaN, quiet).
__host__ __device__ inline uint16_t bf16_rne(float x) {
  uint32_t u = f32_bits(x);
  if((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) { return bits_f32((uint32_t)h << 16); }

// OCP e4m3 (E4M3FN: bias 7, max 448, no infinities, S.1111.111 = NaN),
// round-to-nearest-even, saturating to +-448.
__host__ __device__ inline uint8_t e4m3_rne_sat(float x) {
  uint32_t u = f32_bits(x);
  uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
  uint32_t au = u & 0x7fffffffu;
  if(au > 0x7f800000u) return (uint8_t)(sign | 0x7f);
  float a = bits_f32(au);
  if(a > 448.0f) return (uint8_t)(sign | 0x7e);
  if(a < 0.015625f) {  // below the smallest normal 2^-6: multiples of 2^-9
    float s = a * 512.0f;  // exact
    // round half to even on an exact value in [0, 8)
    uint32_t i = (uint32_t)s;
    float frac = s - (float)i;
    if(frac > 0.5f || (frac == 0.5f && (i & 1u))) ++i;
    return (uint8_t)(sign | i);  // i == 8 encodes the smallest normal
  }
  int e = (int)((au >> 23) & 0xffu) - 127;  // -6 .. 8
  uint32_t m = au & 0x7fffffu;
  uint32_t r = (m + 0x7ffffu + ((m >> 20) & 1u)) >> 20;  // 0 .. 8
  uint32_t code = ((uint32_t)(e + 7) << 3) + r;
  if(code > 0x7eu) code = 0x7eu;
  return (uint8_t)(sign | code);
}
__host__ __device__ inline float e4m3_to_f32(uint8_t q) {
  uint32_t s = q & 0x80u, e = (q >> 3) & 0xfu, m = q & 7u;
  float v;
  if(e == 15 && m == 7) return bits_f32(0x7fc00000u | (s << 24));
  if(e == 0)
    v = (float)m * 0.001953125f;  // m * 2^-9
  else
    v = bits_f32(((e + 120u) << 23) | (m << 20));
  return s ? -v : v;
}

// Smallest e with amax <= 448 * 2^e (0 for amax == 0), clamped to the E8M0
// range usable as an MFMA scale.
__host__ __device__ inline int fp8_block_exponent(float amax) {
  if(!(amax > 0.0f)) return 0;
  int e = (int)((f32_bits(amax) >> 23) & 0xffu) - 127 - 8;
  if(((f32_bits(amax) >> 23) & 0xffu) == 0) e = -126 - 8;
  // adjust so that 448*2^(e-1) < amax <= 448*2^e, using exact power-of-two scaling
  for(int it = 0; it < 4; ++it) {
    if(amax > ldexpf(448.0f, e)) ++e;
    else if(amax <= ldexpf(448.0f, e - 1)) --e;
    else break;
  }
  if(e < -127) e = -127;
  if(e > 127) e = 127;
  return e;
}

// Keyed SplitMix64: the value for (seed, matrix, index) is independent of
// launch shape, so host and device generate bit-identical inputs.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t keyed_hash(uint64_t seed, uint32_t matrix, uint64_t index) {
  uint64_t base = mix64(seed ^ (0x9e3779b97f4a7c15ull * ((uint64_t)matrix + 1ull)));
  return mix64(base + 0x9e3779b97f4a7c15ull * (index + 1ull));
}
// 2 * (h >> 40) * 2^-24 - 1: a 24-bit grid on [-1, 1), exact in FP32.
__host__ __device__ inline float uniform_pm1(uint64_t h) {
  return (float)(uint32_t)(h >> 40) * 1.1920928955078125e-07f - 1.0f;
}

// ---------------------------------------------------------------------------
// Tile scheduling
// ---------------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs (wg i -> XCD i % 8).
// This renumbers them so each XCD owns one contiguous range of tile ids, which
// keeps the A/B panels a tile needs in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int xcd = bid & 7, local = bid >> 3;
  const int per = nblocks >> 3, rem = nblocks & 7;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Grouped ordering: GROUP consecutive tile rows are walked column by column.
__device__ __forceinline__ void grouped_tile(int pid, int tiles_m, int tiles_n, int group, int& tm, int& tn) {
  const int in_group = group * tiles_n;
  const int gid = pid / in_group;
  const int first_m = gid * group;
  const int gsize = min(tiles_m - first_m, group);
  const int r = pid - gid * in_group;
  tm = first_m + r % gsize;
  tn = r / gsize;
}

// The lower triangle of a T x T tile grid in bands of `group` tile rows, each
// band walked column by column (its full-width part, then its own triangle),
// so consecutive ids -- one XCD's workgroups after xcd_remap -- cover a
// compact block of tiles that shares A and B panels in L2, as grouped_tile
// does for a full grid.  A plain row-by-row order gives each round of an
// XCD's workgroups one or two tile rows and as many distinct B panels as
// tiles.  Every (ti, tj) with tj <= ti < T once for
// p < T (T + 1) / 2.
__host__ __device__ inline void grouped_lower_tile(int p, int T, int group, int& ti, int& tj) {
  int r0 = 0, rows = group < T ? group : T;
  for(;;) {
    const int in_band = rows * r0 + rows * (rows + 1) / 2;
    if(p < in_band || r0 + rows >= T) break;
    p -= in_band;
    r0 += rows;
    rows = group < T - r0 ? group : T - r0;
  }
  if(p < rows * r0) {
    tj = p / rows;
    ti = r0 + p % rows;
    return;
  }
  p -= rows * r0;
  int c = 0;
  while(p >= rows - c) {
    p -= rows - c;
    ++c;
  }
  tj = r0 + c;
  ti = r0 + c + p;
}

// A triangular launch's tile (ti, tj) for id p: the triangle order of the
// T x T grid, or -- when `tail_first` >= 0, the re-tiled last round of a
// bigger tile's triangle (tri_tail) -- sub-tile p % r^2 of big tile
// tail_first + p / r^2 of the big_T x big_T triangle.  False for a sub-tile
// with nothing of the triangle in it (above a diagonal big tile's diagonal,
// or past the matrix), which the workgroup skips.
__host__ __device__ inline bool triangle_tile(int p, int T, int group, int tail_first, int r, int big_T, int& ti,
    int& tj) {
  if(tail_first < 0) {
    grouped_lower_tile(p, T, group, ti, tj);
    return true;
  }
  int bi, bj;
  grouped_lower_tile(tail_first + p / (r * r), big_T, group, bi, bj);
  const int s = p % (r * r);
  ti = bi * r + s / r;
  tj = bj * r + s % r;
  return tj <= ti && ti < T;
}

// Scheduling modes shared by the FP32 and BF16 kernel families.
enum TileMode : int { TILES_FULL = 0, TILES_LOWER = 1, TILES_UPPER = 2 };

// Epilogue parameters common to all families.
struct Epilogue {
  float* C;
  int64_t ldc;
  float alpha, beta;
  int M, N;
  int uplo;    // TILES_FULL: write everything; LOWER/UPPER: that triangle only
  int mirror;  // also write the opposite triangle from the transposed value
};

typedef float floatx16 __attribute__((ext_vector_type(16)));

// One 16-byte chunk per lane from global memory straight into LDS
// (global_load_lds_dwordx4): lane l's chunk lands at lds_wave_base + l.  The
// LDS base must be the same for the whole wave.
template <class T>
__device__ __forceinline__ void lds_dma16(const T* src, T* lds_wave_base) {
  static_assert(sizeof(T) == 16, "one 16-byte chunk per lane");
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
      (__attribute__((address_space(3))) void*)lds_wave_base, 16, 0, 0);
}

// Waits for this wave's outstanding memory operations, LDS transfers included.
__device__ __forceinline__ void lds_dma_wait() { __builtin_amdgcn_s_waitcnt(0); }
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Compile-time loop: f(std::integral_constant<int, I>) for I in [0, N).  Used
// where an index selects an accumulator block, so the index stays a constant
// and the accumulators stay in registers (a runtime index, or a loop the
// unroller gives up on, moves the whole array to scratch memory).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Writes one 32x32 MFMA accumulator block whose top-left corner is (r0, c0).
// v_mfma_f32_32x32x* layout: lane l, register q holds
// row 8*(q/4) + 4*(l/32) + q%4, column l%32.
// Always inlined and taking the block by value: a call would force the whole
// accumulator array into scratch memory.
__device__ __forceinline__ void store_block32(const Epilogue& ep, floatx16 acc, int r0, int c0, int lane) {
  const int col = c0 + (lane & 31);
  const int rbase = r0 + 4 * (lane >> 5);
  if(ep.uplo == TILES_FULL && r0 + 32 <= ep.M && c0 + 32 <= ep.N) {
    // Interior block: each register q is one 32-float row segment per half wave.
    float* c = ep.C + (int64_t)rbase * ep.ldc + col;
    if(ep.beta != 0.0f) {
      float old[16];
#pragma unroll
      for(int q = 0; q < 16; ++q) old[q] = c[(int64_t)(8 * (q >> 2) + (q & 3)) * ep.ldc];
#pragma unroll
      for(int q = 0; q < 16; ++q) c[(int64_t)(8 * (q >> 2) + (q & 3)) * ep.ldc] = ep.alpha * acc[q] + ep.beta * old[q];
    } else {
#pragma unroll
      for(int q = 0; q < 16; ++q) c[(int64_t)(8 * (q >> 2) + (q & 3)) * ep.ldc] = ep.alpha * acc[q];
    }
    return;
  }
#pragma unroll
  for(int q = 0; q < 16; ++q) {
    const int row = rbase + 8 * (q >> 2) + (q & 3);
    if(row >= ep.M || col >= ep.N) continue;
    bool direct = true;
    if(ep.uplo == TILES_LOWER) direct = row >= col;
    else if(ep.uplo == TILES_UPPER) direct = row <= col;
    if(direct) {
      float* c = ep.C + (int64_t)row * ep.ldc + col;
      float v = ep.alpha * acc[q];
      if(ep.beta != 0.0f) v += ep.beta * *c;
      *c = v;
    }
    if(ep.mirror && ep.uplo != TILES_FULL && row != col && direct) {
      float* c = ep.C + (int64_t)col * ep.ldc + row;
      float v = ep.alpha * acc[q];
      if(ep.beta != 0.0f) v += ep.beta * *c;
      *c = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers implemented by the kernel translation units
// ---------------------------------------------------------------------------
struct DenseOp {
  // C = alpha * sum_p opA_p * opB_p + beta * C, with p over one or two products
  // (two for syr2k: A*B^T then B*A^T).
  int M, N, K;
  const float* A; int64_t lda; bool a_kmajor;  // element (m,k) at A[m*lda+k] if kmajor else A[k*lda+m]
  const float* B; int64_t ldb; bool b_kmajor;  // element (k,n) at B[n*ldb+k] if kmajor else B[k*ldb+n]
  int nprod;                                    // 1, or 2 with operands swapped for the second
  Epilogue ep;
  int tiles;                                    // TILES_FULL / LOWER / UPPER
};

// Triangular schedules whose last round of tiles leaves most CUs idle: the
// full rounds run on the chosen square tile, and the big tiles of the last
// round are re-tiled `ratio` x `ratio` into a smaller square configuration
// in a second launch, when that is estimated faster than one more round of
// big tiles (small == -1: no split).
struct TriTail {
  int head = 0;    // big tiles of the first launch (all of them without a split)
  int small = -1;  // configuration of the second launch
  int ratio = 1;   // big tile side / small tile side
};
TriTail tri_tail(mfx_precision p, int config, int64_t tiles_per_side, int cus);

mfx_status launch_dense_f32(Handle* h, const DenseOp& op, int config, hipStream_t s);
// The same with split-K partial tiles in `ws` (ws_bytes long) instead of the
// handle's workspace; fails if a split this call makes needs more.
mfx_status launch_dense_f32_ws(Handle* h, const DenseOp& op, int config, void* ws, size_t ws_bytes, hipStream_t s);
// FP32 split-K pieces (gemm_f32.hip): k slices per tile for a configuration
// (1: none), the partial buffer's size, and the launch on given buffers.
int f32_splits(int config, int64_t tiles, int64_t K, int cus);
size_t f32_splitk_workspace(int config, int64_t tiles, int splits);
mfx_status launch_f32_splitk(Handle* h, const DenseOp& op, int config, int splits, float* partial, hipStream_t s);
mfx_status launch_dense_bf16x(Handle* h, const DenseOp& op, bool split, int config, hipStream_t s);
// The same on a caller-provided workspace of dense_bf16x_workspace() bytes.
mfx_status launch_dense_bf16x_ws(Handle* h, const DenseOp& op, bool split, int config, void* ws, hipStream_t s);
mfx_status launch_dense_fp8(Handle* h, const DenseOp& op, bool two_term, int config, hipStream_t s);
size_t dense_bf16x_workspace(const DenseOp& op, bool split, int config);
size_t dense_fp8_workspace(const DenseOp& op, bool two_term, int config);

int config_count(mfx_precision p);
bool config_info(mfx_precision p, int c, int* bm, int* bn, int* bk, int* waves);
int choose_config(Handle* h, mfx_precision p, int64_t m, int64_t n, int64_t k);

}  // namespace mfx

struct mfx_handle_s {
  mfx::Handle h;
};
