// Split-BF16 and single-BF16 dense contractions on v_mfma_f32_32x32x16_bf16.
//
// Split BF16: every FP32 operand x is written as hi + lo with
// hi = bf16_rne(x), lo = bf16_rne(x - hi); the product uses
// lo*hi + hi*lo + hi*hi (the lo*lo term, ~2^-16 relative, is dropped) with
// FP32 accumulation, which keeps the relative Frobenius error near 4e-6 for
// uniform inputs.  Single BF16 keeps only hi*hi.
//
// Two kernels per call:
//  1. pack: converts each operand once into "fragment order"
//       P[blk][kb][hl][lane][8]   (bf16)
//     where blk is a 32-row block, kb a 16-wide k block, hl = hi / lo, and lane
//     l holds row blk*32 + l%32, k = kb*16 + 8*(l/32) + 0..7 -- exactly the
//     register operand of the 32x32x16 MFMA.  Rows are padded to the tile
//     size and k to the k-step with zeros, so the main loop has no bounds
//     checks and every load is a contiguous 16-byte chunk.
//  2. contract: BM x BN tiles, BK = 32 (two MFMA k-steps) per stage, double-
//     buffered LDS in the same fragment order (ds_read_b128 of consecutive
//     16-byte chunks: conflict-free), fused alpha/beta epilogue.
// The same packed layout serves an operand in either role (A rows, or B
// columns), which is what lets syrk use one packed copy of A for both.
#include "mfx_internal.h"

namespace mfx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// Native clang vector (not HIP's uint4 struct) so register arrays of chunks are
// promoted to registers instead of living in scratch memory.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int KB_PER_STAGE = 2;  // BK = 32

// ---------------------------------------------------------------------------
// Pack
// ---------------------------------------------------------------------------
struct PackParams {
  const float* src;
  int64_t ld;
  int rows, K;       // logical extent of the operand (rows x K)
  int kmajor;        // element (r, k) at src[r*ld + k], else src[k*ld + r]
  int vec;           // 128-bit loads allowed (kmajor only)
  int blocks;        // padded rows / 32
  int kbs;           // padded K / 16
  int nhl;           // 2 = hi and lo, 1 = hi only
  uint4* dst;
};

__device__ inline uint32_t pack2(uint16_t a, uint16_t b) { return (uint32_t)a | ((uint32_t)b << 16); }

__global__ __launch_bounds__(256) void pack_bf16x_kernel(const PackParams p) {
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = (int64_t)p.blocks * p.kbs * 64;
  if(gid >= total) return;
  const int lane = (int)(gid & 63);
  const int64_t t = gid >> 6;
  const int kb = (int)(t % p.kbs);
  const int blk = (int)(t / p.kbs);
  const int r = blk * 32 + (lane & 31);
  const int k0 = kb * 16 + (lane >> 5) * 8;
  float v[8];
  if(p.kmajor) {
    const float* s = p.src + (int64_t)r * p.ld + k0;
    if(r < p.rows && p.vec && k0 + 8 <= p.K) {
      const float4 x = reinterpret_cast<const float4*>(s)[0];
      const float4 y = reinterpret_cast<const float4*>(s)[1];
      v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
      v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    } else {
#pragma unroll
      for(int e = 0; e < 8; ++e) v[e] = (r < p.rows && k0 + e < p.K) ? s[e] : 0.0f;
    }
  } else {
#pragma unroll
    for(int e = 0; e < 8; ++e) v[e] = (r < p.rows && k0 + e < p.K) ? p.src[(int64_t)(k0 + e) * p.ld + r] : 0.0f;
  }
  uint16_t hi[8], lo[8];
#pragma unroll
  for(int e = 0; e < 8; ++e) {
    hi[e] = bf16_rne(v[e]);
    lo[e] = bf16_rne(v[e] - bf16_to_f32(hi[e]));
  }
  const int64_t base = ((int64_t)blk * p.kbs + kb) * p.nhl * 64 + lane;
  p.dst[base] = make_uint4(pack2(hi[0], hi[1]), pack2(hi[2], hi[3]), pack2(hi[4], hi[5]), pack2(hi[6], hi[7]));
  if(p.nhl == 2)
    p.dst[base + 64] = make_uint4(pack2(lo[0], lo[1]), pack2(lo[2], lo[3]), pack2(lo[4], lo[5]), pack2(lo[6], lo[7]));
}

// ---------------------------------------------------------------------------
// Contract
// ---------------------------------------------------------------------------
struct BfParams {
  const u32x4* PA;  // packed A-role operand of product 0
  const u32x4* PB;  // packed B-role operand of product 0
  int kbs;          // k blocks of 16 in both packed operands
  int nprod;        // 2: second product uses (PB, PA)
  int tiles_m, tiles_n, group;
  int mode;         // TILES_FULL / LOWER / UPPER
  Epilogue ep;
};

template <int BM, int BN, int WM, int WN, int NHL>
struct BfCfg {
  static constexpr int NWM = BM / WM, NWN = BN / WN, NW = NWM * NWN, NT = NW * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int CH_A = BM / 32 * KB_PER_STAGE * NHL * 64;  // uint4 chunks per stage
  static constexpr int CH_B = BN / 32 * KB_PER_STAGE * NHL * 64;
  static constexpr int PER_A = CH_A / NT, PER_B = CH_B / NT;
  static_assert(CH_A % NT == 0 && CH_B % NT == 0, "stage must split evenly");
};

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 u) { return __builtin_bit_cast(bf16x8, u); }

template <int BM, int BN, int WM, int WN, int NHL>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_bf16x_kernel(const BfParams p) {
  using Cfg = BfCfg<BM, BN, WM, WN, NHL>;
  constexpr int NT = Cfg::NT, FM = Cfg::FM, FN = Cfg::FN;
  constexpr int SA = Cfg::CH_A, SB = Cfg::CH_B;
  __shared__ u32x4 lds[2 * (SA + SB)];

  int tm, tn;
  const int pid = xcd_remap(blockIdx.x, gridDim.x);
  if(p.mode == TILES_FULL) {
    grouped_tile(pid, p.tiles_m, p.tiles_n, p.group, tm, tn);
  } else {
    int ti, tj;
    lower_tile(pid, ti, tj);
    if(p.mode == TILES_LOWER) { tm = ti; tn = tj; } else { tm = tj; tn = ti; }
  }
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid / Cfg::NWN, wn = wid % Cfg::NWN;

  floatx16 acc[FM][FN];
#pragma unroll
  for(int i = 0; i < FM; ++i)
#pragma unroll
    for(int j = 0; j < FN; ++j)
#pragma unroll
      for(int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  // Chunk c of a stage: block c / (KB_PER_STAGE*NHL*64) of the tile, then the
  // contiguous (kb, hl, lane) run of that block for this stage.
  constexpr int RUN = KB_PER_STAGE * NHL * 64;
  const int KT = p.kbs / KB_PER_STAGE;
  const int steps = KT * p.nprod;
  const int64_t blk_stride = (int64_t)p.kbs * NHL * 64;
  u32x4 ra[Cfg::PER_A], rb[Cfg::PER_B];

  auto issue = [&](int s) __attribute__((always_inline)) {
    const int prod = s / KT, kt = s - prod * KT;
    const u32x4* A = prod ? p.PB : p.PA;
    const u32x4* B = prod ? p.PA : p.PB;
    const int64_t koff = (int64_t)kt * RUN;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i) {
      const int c = tid + i * NT;
      const int blk = c / RUN, run = c - blk * RUN;
      ra[i] = A[(int64_t)(tm * (BM / 32) + blk) * blk_stride + koff + run];
    }
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i) {
      const int c = tid + i * NT;
      const int blk = c / RUN, run = c - blk * RUN;
      rb[i] = B[(int64_t)(tn * (BN / 32) + blk) * blk_stride + koff + run];
    }
  };
  auto commit = [&](int buf) __attribute__((always_inline)) {
    u32x4* sa = lds + buf * (SA + SB);
    u32x4* sb = sa + SA;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i) sa[tid + i * NT] = ra[i];
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i) sb[tid + i * NT] = rb[i];
  };
  auto compute = [&](int cur) __attribute__((always_inline)) {
    const u32x4* sa = lds + cur * (SA + SB);
    const u32x4* sb = sa + SA;
#pragma unroll
    for(int kb = 0; kb < KB_PER_STAGE; ++kb) {
      bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for(int i = 0; i < FM; ++i) {
        const int blk = wm * FM + i;
        ah[i] = as_bf16x8(sa[(blk * KB_PER_STAGE + kb) * NHL * 64 + lane]);
        if(NHL == 2) al[i] = as_bf16x8(sa[((blk * KB_PER_STAGE + kb) * NHL + 1) * 64 + lane]);
      }
#pragma unroll
      for(int j = 0; j < FN; ++j) {
        const int blk = wn * FN + j;
        bh[j] = as_bf16x8(sb[(blk * KB_PER_STAGE + kb) * NHL * 64 + lane]);
        if(NHL == 2) bl[j] = as_bf16x8(sb[((blk * KB_PER_STAGE + kb) * NHL + 1) * 64 + lane]);
      }
      // Product-major order: consecutive MFMAs write different accumulators,
      // so none waits on the previous one's result.  The small cross terms go
      // in first so they are added while the accumulator is smallest.
      if(NHL == 2) {
#pragma unroll
        for(int i = 0; i < FM; ++i)
#pragma unroll
          for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for(int i = 0; i < FM; ++i)
#pragma unroll
          for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
      }
#pragma unroll
      for(int i = 0; i < FM; ++i)
#pragma unroll
        for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    }
  };

  // Steady state without branches around the prefetch, so the registers
  // holding stage s+1 stay live across the MFMAs of stage s (no early wait).
  if(steps > 0) {
    issue(0);
    commit(0);
  }
  __syncthreads();
  for(int s = 0; s + 1 < steps; ++s) {
    issue(s + 1);
    compute(s & 1);
    commit((s + 1) & 1);
    __syncthreads();
  }
  if(steps > 0) compute((steps - 1) & 1);

  static_for<FM * FN>([&](auto ij) __attribute__((always_inline)) {
    constexpr int i = decltype(ij)::value / FN, j = decltype(ij)::value % FN;
    store_block32(p.ep, acc[i][j], tm * BM + wm * WM + i * 32, tn * BN + wn * WN + j * 32, lane);
  });
}

struct BfConfig {
  int bm, bn, waves;
};
constexpr BfConfig kBfConfigs[] = {
    {256, 256, 4},  // 0: 128x128 per wave
    {128, 128, 4},  // 1: 64x64 per wave
    {64, 64, 4},    // 2: 32x32 per wave
    {256, 128, 4},  // 3: 128x64 per wave
};
constexpr int kBfConfigCount = sizeof(kBfConfigs) / sizeof(kBfConfigs[0]);

template <int BM, int BN, int WM, int WN>
mfx_status launch_bf_cfg(Handle* h, const BfParams& p, bool split, int grid, hipStream_t s) {
  if(split)
    return timed_launch(h, "gemm_split_bf16_mfma", s, [&] {
      gemm_bf16x_kernel<BM, BN, WM, WN, 2><<<grid, BfCfg<BM, BN, WM, WN, 2>::NT, 0, s>>>(p);
    });
  return timed_launch(h, "gemm_bf16_mfma", s, [&] {
    gemm_bf16x_kernel<BM, BN, WM, WN, 1><<<grid, BfCfg<BM, BN, WM, WN, 1>::NT, 0, s>>>(p);
  });
}

struct PackPlan {
  int64_t rows_pad_a, rows_pad_b, kbs;
  size_t bytes_a, bytes_b;
};

PackPlan plan(const DenseOp& op, bool split, int config) {
  const BfConfig& c = kBfConfigs[config];
  PackPlan pl;
  const int nhl = split ? 2 : 1;
  pl.rows_pad_a = round_up(op.M, c.bm);
  pl.rows_pad_b = round_up(op.N, c.bn);
  pl.kbs = round_up(op.K, 16 * KB_PER_STAGE) / 16;
  pl.bytes_a = (size_t)pl.rows_pad_a * pl.kbs * 16 * nhl * 2;
  pl.bytes_b = (size_t)pl.rows_pad_b * pl.kbs * 16 * nhl * 2;
  return pl;
}

mfx_status pack(Handle* h, const float* src, int64_t ld, int rows, int K, bool kmajor, int64_t rows_pad, int64_t kbs,
    int nhl, uint4* dst, hipStream_t s) {
  PackParams pp;
  pp.src = src;
  pp.ld = ld;
  pp.rows = rows;
  pp.K = K;
  pp.kmajor = kmajor;
  pp.vec = kmajor && (reinterpret_cast<uintptr_t>(src) & 15u) == 0 && ld % 4 == 0;
  pp.blocks = (int)(rows_pad / 32);
  pp.kbs = (int)kbs;
  pp.nhl = nhl;
  pp.dst = dst;
  const int64_t total = rows_pad / 32 * kbs * 64;
  if(total == 0) return MFX_SUCCESS;
  const int grid = (int)ceil_div(total, 256);
  return timed_launch(h, nhl == 2 ? "pack_split_bf16" : "pack_bf16", s,
      [&] { pack_bf16x_kernel<<<grid, 256, 0, s>>>(pp); });
}

}  // namespace

int bf_config_count() { return kBfConfigCount; }
bool bf_config_info(int c, int* bm, int* bn, int* bk, int* waves) {
  if(c < 0 || c >= kBfConfigCount) return false;
  *bm = kBfConfigs[c].bm;
  *bn = kBfConfigs[c].bn;
  *bk = 16 * KB_PER_STAGE;
  *waves = kBfConfigs[c].waves;
  return true;
}

size_t dense_bf16x_workspace(const DenseOp& op, bool split, int config) {
  if(config < 0 || config >= kBfConfigCount) return 0;
  PackPlan pl = plan(op, split, config);
  // syrk (single operand used in both roles) still reserves two slots; the
  // second is simply unused.
  return round_up(pl.bytes_a, 256) + round_up(pl.bytes_b, 256);
}

mfx_status launch_dense_bf16x(Handle* h, const DenseOp& op, bool split, int config, hipStream_t s) {
  if(config < 0 || config >= kBfConfigCount) return invalid("BF16 tile configuration out of range");
  const BfConfig& c = kBfConfigs[config];
  if(op.M == 0 || op.N == 0) return MFX_SUCCESS;
  const int nhl = split ? 2 : 1;
  PackPlan pl = plan(op, split, config);
  const bool same = op.A == op.B && op.lda == op.ldb && op.a_kmajor == op.b_kmajor && op.M == op.N &&
                    pl.rows_pad_a == pl.rows_pad_b;
  void* ws = nullptr;
  MFX_CHECK(workspace(h, dense_bf16x_workspace(op, split, config), &ws));
  uint4* pa = static_cast<uint4*>(ws);
  uint4* pb = reinterpret_cast<uint4*>(static_cast<char*>(ws) + round_up(pl.bytes_a, 256));
  // The B-role operand is indexed (n, k): B[n*ldb + k] if b_kmajor else B[k*ldb + n].
  MFX_CHECK(pack(h, op.A, op.lda, op.M, op.K, op.a_kmajor, pl.rows_pad_a, pl.kbs, nhl, pa, s));
  if(same) {
    pb = pa;
  } else {
    MFX_CHECK(pack(h, op.B, op.ldb, op.N, op.K, op.b_kmajor, pl.rows_pad_b, pl.kbs, nhl, pb, s));
  }
  BfParams p;
  p.PA = reinterpret_cast<const u32x4*>(pa);
  p.PB = reinterpret_cast<const u32x4*>(pb);
  p.kbs = (int)pl.kbs;
  p.nprod = op.nprod;
  p.tiles_m = (int)(pl.rows_pad_a / c.bm);
  p.tiles_n = (int)(pl.rows_pad_b / c.bn);
  p.group = 8;
  p.mode = op.tiles;
  p.ep = op.ep;
  int grid;
  if(op.tiles == TILES_FULL) {
    grid = p.tiles_m * p.tiles_n;
  } else {
    if(p.tiles_m != p.tiles_n || c.bm != c.bn) return invalid("triangular schedule needs square tiles");
    grid = p.tiles_m * (p.tiles_m + 1) / 2;
  }
  if(op.nprod == 2 && pl.rows_pad_a != pl.rows_pad_b) return invalid("two-product form needs equal operand shapes");
  switch(config) {
    case 0: return launch_bf_cfg<256, 256, 128, 128>(h, p, split, grid, s);
    case 1: return launch_bf_cfg<128, 128, 64, 64>(h, p, split, grid, s);
    case 2: return launch_bf_cfg<64, 64, 32, 32>(h, p, split, grid, s);
    case 3: return launch_bf_cfg<256, 128, 128, 64>(h, p, split, grid, s);
  }
  return invalid("unknown BF16 configuration");
}

// Element-wise split used by mfx_split_bf16 (a test and tooling utility).
namespace {
__global__ void split_elementwise_kernel(int64_t n, const float* x, uint16_t* hi, uint16_t* lo) {
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint16_t h = bf16_rne(x[i]);
    hi[i] = h;
    lo[i] = bf16_rne(x[i] - bf16_to_f32(h));
  }
}
}  // namespace

mfx_status launch_split_elementwise(Handle* h, int64_t n, const float* x, uint16_t* hi, uint16_t* lo, hipStream_t s) {
  if(n == 0) return MFX_SUCCESS;
  const int grid = (int)std::min<int64_t>(ceil_div(n, 256), 65536);
  return timed_launch(h, "split_bf16_elementwise", s, [&] { split_elementwise_kernel<<<grid, 256, 0, s>>>(n, x, hi, lo); });
}

}  // namespace mfx
