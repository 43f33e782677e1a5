// Split-BF16 and single-BF16 dense contractions on v_mfma_f32_32x32x16_bf16
// (building blocks in tile_bf16.h).
//
// Two kernels per call:
//  1. pack: converts each operand once into fragment order (hi and lo), with
//     zero padding, so the contraction loop has no bounds checks and every
//     load is a contiguous 16-byte chunk;
//  2. contract: BM x BN tiles, BK = 32 (two MFMA k-steps) per stage,
//     double-buffered LDS in the same fragment order (ds_read_b128 of
//     consecutive 16-byte chunks: conflict-free), fused alpha/beta epilogue.
// The same packed layout serves an operand in either role (A rows, or B
// columns), which is what lets syrk use one packed copy of A for both.
#include "tile_bf16.h"

namespace mfx {
namespace {

struct PackParams {
  PackSrc s;
  int blocks;  // padded rows / 32
  int kbs;     // padded K / 16
  u32x4* dst;
};

// One lane-chunk per thread; a non-k-major source goes through the wave's
// LDS tile (load_chunk8_staged) so its reads are 128-byte row segments, with
// the (blk, kb) pairs walked kb-major so that neighbouring waves read
// neighbouring segments of the same source rows.
template <int NHL>
__global__ __launch_bounds__(256) void pack_bf16x_kernel(const PackParams p) {
  __shared__ float tiles[4][STAGED_TILE_FLOATS];
  const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = (int64_t)p.blocks * p.kbs * 64;
  const bool active = gid < total;
  const int lane = (int)(gid & 63);
  const int64_t t = gid >> 6;
  if(p.s.kmajor) {
    if(active) pack_chunk<NHL>(p.s, (int)(t / p.kbs), (int)(t % p.kbs), lane, p.kbs, p.dst);
    return;
  }
  const int blk = (int)(t % p.blocks), kb = (int)(t / p.blocks);
  float v[8];
  load_chunk8_staged(p.s, blk, kb, lane, active, v, tiles[threadIdx.x >> 6]);
  if(active) store_chunk<NHL>(p.dst, blk, kb, lane, p.kbs, v);
}

struct BfParams {
  BfTileArgs t;
  int tiles_m, tiles_n, group;
  int mode;  // TILES_FULL / LOWER / UPPER
  int tail_first = -1, tail_ratio = 1, big_tiles = 0;  // triangle_tile's re-tiled last round
};

template <int BM, int BN, int WM, int WN, int NHL>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_bf16x_kernel(const BfParams p) {
  __shared__ u32x4 lds[BfCfg<BM, BN, WM, WN, NHL>::SMEM_CHUNKS];
  int tm, tn;
  const int pid = xcd_remap(blockIdx.x, gridDim.x);
  if(p.mode == TILES_FULL) {
    grouped_tile(pid, p.tiles_m, p.tiles_n, p.group, tm, tn);
  } else {
    int ti, tj;
    if(!triangle_tile(pid, p.tiles_m, p.group, p.tail_first, p.tail_ratio, p.big_tiles, ti, tj)) return;
    if(p.mode == TILES_LOWER) { tm = ti; tn = tj; } else { tm = tj; tn = ti; }
  }
  bf16_tile<BM, BN, WM, WN, NHL>(p.t, tm, tn, lds);
}

struct BfConfig {
  int bm, bn, waves;
};
constexpr BfConfig kBfConfigs[] = {
    {256, 256, 4},  // 0: 128x128 per wave
    {128, 128, 4},  // 1: 64x64 per wave
    {64, 64, 4},    // 2: 32x32 per wave
    {256, 128, 4},  // 3: 128x64 per wave
    {256, 256, 8},  // 4: 128x64 per wave, two waves per SIMD
    {256, 128, 8},  // 5: 64x64 per wave, two waves per SIMD
    {128, 128, 8},  // 6: 64x32 per wave, two waves per SIMD
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

mfx_status launch_bf_config(Handle* h, const BfParams& p, bool split, int config, int grid, hipStream_t s) {
  switch(config) {
    case 0: return launch_bf_cfg<256, 256, 128, 128>(h, p, split, grid, s);
    case 1: return launch_bf_cfg<128, 128, 64, 64>(h, p, split, grid, s);
    case 2: return launch_bf_cfg<64, 64, 32, 32>(h, p, split, grid, s);
    case 3: return launch_bf_cfg<256, 128, 128, 64>(h, p, split, grid, s);
    case 4: return launch_bf_cfg<256, 256, 128, 64>(h, p, split, grid, s);
    case 5: return launch_bf_cfg<256, 128, 64, 64>(h, p, split, grid, s);
    case 6: return launch_bf_cfg<128, 128, 64, 32>(h, p, split, grid, s);
  }
  return invalid("unknown BF16 configuration");
}

struct PackPlan {
  int64_t rows_pad_a, rows_pad_b, kbs;
  size_t bytes_a, bytes_b;
};

PackPlan plan(const DenseOp& op, bool split, int config) {
  const BfConfig& c = kBfConfigs[config];
  PackPlan pl;
  const int nhl = split ? 2 : 1;
  // rows padded to 32 only (bf16_tile reads the last block again past them)
  pl.rows_pad_a = round_up(op.M, 32);
  pl.rows_pad_b = round_up(op.N, 32);
  (void)c;
  pl.kbs = round_up(op.K, 16 * kb_per_stage_of(nhl)) / 16;
  pl.bytes_a = (size_t)pl.rows_pad_a * pl.kbs * 16 * nhl * 2;
  pl.bytes_b = (size_t)pl.rows_pad_b * pl.kbs * 16 * nhl * 2;
  return pl;
}

mfx_status pack(Handle* h, const float* src, int64_t ld, int rows, int K, bool kmajor, int64_t rows_pad, int64_t kbs,
    int nhl, u32x4* dst, hipStream_t s) {
  PackParams pp;
  pp.s = make_pack_src(src, ld, rows, K, kmajor);
  pp.blocks = (int)(rows_pad / 32);
  pp.kbs = (int)kbs;
  pp.dst = dst;
  const int64_t total = rows_pad / 32 * kbs * 64;
  if(total == 0) return MFX_SUCCESS;
  const int grid = (int)ceil_div(total, 256);
  if(nhl == 2)
    return timed_launch(h, "pack_split_bf16", s, [&] { pack_bf16x_kernel<2><<<grid, 256, 0, s>>>(pp); });
  return timed_launch(h, "pack_bf16", s, [&] { pack_bf16x_kernel<1><<<grid, 256, 0, s>>>(pp); });
}

}  // namespace

PackSrc make_pack_src(const float* src, int64_t ld, int rows, int K, bool kmajor) {
  PackSrc s;
  s.src = src;
  s.ld = ld;
  s.rows = rows;
  s.K = K;
  s.kmajor = kmajor;
  // 128-bit loads along the contiguous dimension (k if kmajor, rows if not)
  s.vec = (reinterpret_cast<uintptr_t>(src) & 15u) == 0 && ld % 4 == 0;
  return s;
}

int bf_config_count() { return kBfConfigCount; }
bool bf_config_info(int c, bool split, int* bm, int* bn, int* bk, int* waves) {
  if(c < 0 || c >= kBfConfigCount) return false;
  *bm = kBfConfigs[c].bm;
  *bn = kBfConfigs[c].bn;
  *bk = 16 * kb_per_stage_of(split ? 2 : 1);
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
  if(op.M == 0 || op.N == 0) return MFX_SUCCESS;
  void* ws = nullptr;
  MFX_CHECK(workspace(h, dense_bf16x_workspace(op, split, config), &ws));
  return launch_dense_bf16x_ws(h, op, split, config, ws, s);
}

// The pack kernels write every chunk of both padded panels (zeros past the
// matrix), so `ws` needs no clearing and may hold anything on entry.
mfx_status launch_dense_bf16x_ws(Handle* h, const DenseOp& op, bool split, int config, void* ws, hipStream_t s) {
  if(config < 0 || config >= kBfConfigCount) return invalid("BF16 tile configuration out of range");
  const BfConfig& c = kBfConfigs[config];
  if(op.M == 0 || op.N == 0) return MFX_SUCCESS;
  const int nhl = split ? 2 : 1;
  PackPlan pl = plan(op, split, config);
  const bool same = op.A == op.B && op.lda == op.ldb && op.a_kmajor == op.b_kmajor && op.M == op.N &&
                    pl.rows_pad_a == pl.rows_pad_b;
  if(op.nprod == 2 && pl.rows_pad_a != pl.rows_pad_b) return invalid("two-product form needs equal operand shapes");
  u32x4* pa = static_cast<u32x4*>(ws);
  u32x4* pb = reinterpret_cast<u32x4*>(static_cast<char*>(ws) + round_up(pl.bytes_a, 256));
  // The B-role operand is indexed (n, k): B[n*ldb + k] if b_kmajor else B[k*ldb + n].
  MFX_CHECK(pack(h, op.A, op.lda, op.M, op.K, op.a_kmajor, pl.rows_pad_a, pl.kbs, nhl, pa, s));
  if(same) {
    pb = pa;
  } else {
    MFX_CHECK(pack(h, op.B, op.ldb, op.N, op.K, op.b_kmajor, pl.rows_pad_b, pl.kbs, nhl, pb, s));
  }
  BfParams p;
  p.t.PA = pa;
  p.t.PB = pb;
  p.t.kbs = (int)pl.kbs;
  p.t.nprod = op.nprod;
  p.t.ep = op.ep;
  p.t.nblk_a = (int)(pl.rows_pad_a / 32);
  p.t.nblk_b = (int)(pl.rows_pad_b / 32);
  p.tiles_m = (int)ceil_div(op.M, c.bm);
  p.tiles_n = (int)ceil_div(op.N, c.bn);
  p.group = 8;
  p.mode = op.tiles;
  int grid;
  if(op.tiles == TILES_FULL) {
    grid = p.tiles_m * p.tiles_n;
  } else {
    if(p.tiles_m != p.tiles_n || c.bm != c.bn) return invalid("triangular schedule needs square tiles");
    grid = p.tiles_m * (p.tiles_m + 1) / 2;
  }
  if(grid == 0) return MFX_SUCCESS;
  if(op.tiles != TILES_FULL) {
    const TriTail tail =
        tri_tail(split ? MFX_PREC_SPLIT_BF16 : MFX_PREC_BF16, config, p.tiles_m, h->cu_count > 0 ? h->cu_count : 256);
    if(tail.small >= 0) {
      // the full rounds on this tile, the last round's big tiles re-tiled
      MFX_CHECK(launch_bf_config(h, p, split, config, tail.head, s));
      BfParams q = p;
      q.tiles_m = q.tiles_n = (int)ceil_div(op.M, kBfConfigs[tail.small].bm);
      q.tail_first = tail.head;
      q.tail_ratio = tail.ratio;
      q.big_tiles = p.tiles_m;
      return launch_bf_config(h, q, split, tail.small, (grid - tail.head) * tail.ratio * tail.ratio, s);
    }
  }
  return launch_bf_config(h, p, split, config, grid, s);
}

// Element-wise split used by mfx_split_bf16 (a test and tooling utility).
namespace {
__global__ void split_elementwise_kernel(int64_t n, const float* x, uint16_t* hi, uint16_t* lo) {
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    // the conversion the contraction kernels use (split_chunk)
    const uint32_t h = cvt_pk_bf16(x[i], 0.0f) & 0xffffu;
    hi[i] = (uint16_t)h;
    lo[i] = (uint16_t)(cvt_pk_bf16(x[i] - bf16_to_f32((uint16_t)h), 0.0f) & 0xffffu);
  }
}
}  // namespace

mfx_status launch_split_elementwise(Handle* h, int64_t n, const float* x, uint16_t* hi, uint16_t* lo, hipStream_t s) {
  if(n == 0) return MFX_SUCCESS;
  const int grid = (int)std::min<int64_t>(ceil_div(n, 256), 65536);
  return timed_launch(h, "split_bf16_elementwise", s, [&] { split_elementwise_kernel<<<grid, 256, 0, s>>>(n, x, hi, lo); });
}

}  // namespace mfx
