// FP32 dense contractions on v_mfma_f32_32x32x2_f32: one tile per workgroup
// (the tile body is in tile_f32.h).
//
// Replaces the naive one-work-item-per-element kernels of
// polybench/gemm.cpp:77-92, polybench/2mm.cpp:94-123, polybench/3mm.cpp:115-158,
// polybench/syrk.cpp:70-84, polybench/syr2k.cpp:73-88 and
// runtime/matmulchain.cpp:15-29.
#include "tile_f32.h"

namespace mfx {
namespace {

struct F32Params {
  DenseOp op;
  int tiles_m, tiles_n;
  int group;
  int vec_a, vec_b;  // 128-bit global loads allowed for A / B
  int tail_first = -1, tail_ratio = 1, big_tiles = 0;  // triangle_tile's re-tiled last round
};

template <int BM, int BN, int BK, int WM, int WN, int KG, bool AK, bool BKM, bool VEC>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * KG * 64) void gemm_f32_kernel(const F32Params p) {
  __shared__ float smem[F32Cfg<BM, BN, BK, WM, WN, KG>::SMEM_FLOATS];
  int tm, tn;
  const int pid = xcd_remap(blockIdx.x, gridDim.x);
  if(p.op.tiles == TILES_FULL) {
    grouped_tile(pid, p.tiles_m, p.tiles_n, p.group, tm, tn);
  } else {
    int ti, tj;
    if(!triangle_tile(pid, p.tiles_m, p.group, p.tail_first, p.tail_ratio, p.big_tiles, ti, tj)) return;
    if(p.op.tiles == TILES_LOWER) { tm = ti; tn = tj; } else { tm = tj; tn = ti; }
  }
  f32_tile<BM, BN, BK, WM, WN, AK, BKM, VEC, KG>(p.op, tm, tn, smem);
}

// Split-K: a grid that gives the CUs too few tiles (a small square problem)
// runs each tile as `splits` k slices in separate workgroups, several per CU,
// in two kernels.  The first writes each slice's partial tile (alpha = 1,
// beta = 0) to the workspace; the second sums the partials of every output
// element in slice order (so the result is the same on every run) and
// applies alpha / beta to C.  The kernel boundary orders the two: no
// workgroup waits for another and no cache maintenance is needed inside
// either kernel.
struct F32SplitParams {
  DenseOp op;       // the whole problem (one product, full output)
  int tiles_m, tiles_n, group;
  int splits;       // k slices per tile
  int kslice;       // k per slice, a multiple of the tile's BK (the last slice takes the rest)
  float* partial;   // [tm * tiles_n + tn][slice][BM * BN] floats
};

template <int BM, int BN, int BK, int WM, int WN, int KG, bool AK, bool BKM, bool VEC>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * KG * 64) void gemm_f32_splitk_kernel(const F32SplitParams p) {
  __shared__ float smem[F32Cfg<BM, BN, BK, WM, WN, KG>::SMEM_FLOATS];
  // the slices of one tile get consecutive ids, so xcd_remap puts them on
  // one XCD, where they share the tile's operand panels in L2
  const int pid = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = pid / p.splits, slice = pid - tile * p.splits;
  int tm, tn;
  grouped_tile(tile, p.tiles_m, p.tiles_n, p.group, tm, tn);
  const DenseOp& op = p.op;
  const int m0 = tm * BM, n0 = tn * BN;
  const int k0 = slice * p.kslice, k1 = slice + 1 == p.splits ? op.K : k0 + p.kslice;
  // the slice as a one-tile problem of its own: operands offset to the tile's
  // rows / columns and the slice's k range, output the partial tile
  DenseOp sub = op;
  sub.A = op.A + (AK ? (int64_t)m0 * op.lda + k0 : (int64_t)k0 * op.lda + m0);
  sub.B = op.B + (BKM ? (int64_t)n0 * op.ldb + k0 : (int64_t)k0 * op.ldb + n0);
  sub.M = min(BM, op.M - m0);
  sub.N = min(BN, op.N - n0);
  sub.K = k1 - k0;
  float* slot = p.partial + ((int64_t)(tm * p.tiles_n + tn) * p.splits + slice) * (BM * BN);
  sub.ep = Epilogue{slot, BN, 1.0f, 0.0f, sub.M, sub.N, TILES_FULL, 0};
  f32_tile<BM, BN, BK, WM, WN, AK, BKM, VEC, KG>(sub, 0, 0, smem);
}

// One thread per four consecutive columns of one output row (BN is a
// multiple of four, so they lie in one tile row): 128-bit loads of the
// partials, which the slice kernel wrote whole (padding included).
template <int BM, int BN>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const F32SplitParams p) {
  const int cols4 = (p.op.N + 3) >> 2;
  const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(id >= (int64_t)p.op.M * cols4) return;
  const int r = (int)(id / cols4), c = (int)(id - (int64_t)r * cols4) * 4;
  const int tm = r / BM, tn = c / BN;
  const float4* src = reinterpret_cast<const float4*>(
      p.partial + (int64_t)(tm * p.tiles_n + tn) * p.splits * (BM * BN) + (r - tm * BM) * BN + (c - tn * BN));
  float4 v = src[0];
  for(int s = 1; s < p.splits; ++s) {
    const float4 w = src[(int64_t)s * (BM * BN / 4)];
    v.x += w.x;
    v.y += w.y;
    v.z += w.z;
    v.w += w.w;
  }
  const Epilogue& ep = p.op.ep;
  float* cp = ep.C + (int64_t)r * ep.ldc + c;
  const float out[4] = {v.x, v.y, v.z, v.w};
  const int n = min(4, p.op.N - c);
  for(int j = 0; j < n; ++j) cp[j] = ep.beta != 0.0f ? ep.alpha * out[j] + ep.beta * cp[j] : ep.alpha * out[j];
}

struct F32Config {
  int bm, bn, bk, waves;
};
// Compiled tile configurations; index = config id.
constexpr F32Config kF32Configs[] = {
    {256, 256, 16, 4},  // 0: 128x128 per wave, large square problems
    {128, 128, 16, 4},  // 1: 64x64 per wave
    {64, 64, 32, 4},    // 2: 32x32 per wave, small problems
    {256, 128, 16, 4},  // 3: 128x64 per wave
    {256, 256, 16, 8},  // 4: 128x64 per wave, two waves per SIMD
    {256, 128, 16, 8},  // 5: 64x64 per wave, two waves per SIMD
    {128, 128, 16, 8},  // 6: 64x32 per wave, two waves per SIMD
    {64, 64, 64, 16},   // 7: 32x32 per wave, four k-groups of four waves (at most one tile per CU)
};
constexpr int kF32ConfigCount = sizeof(kF32Configs) / sizeof(kF32Configs[0]);

template <int BM, int BN, int BK, int WM, int WN, int KG, bool VEC>
void launch_layout(const F32Params& p, int grid, hipStream_t s) {
  constexpr int NT = F32Cfg<BM, BN, BK, WM, WN, KG>::NT;
  const bool ak = p.op.a_kmajor, bk = p.op.b_kmajor;
  if(ak && !bk) gemm_f32_kernel<BM, BN, BK, WM, WN, KG, true, false, VEC><<<grid, NT, 0, s>>>(p);
  else if(ak && bk) gemm_f32_kernel<BM, BN, BK, WM, WN, KG, true, true, VEC><<<grid, NT, 0, s>>>(p);
  else if(!ak && !bk) gemm_f32_kernel<BM, BN, BK, WM, WN, KG, false, false, VEC><<<grid, NT, 0, s>>>(p);
  else gemm_f32_kernel<BM, BN, BK, WM, WN, KG, false, true, VEC><<<grid, NT, 0, s>>>(p);
}

template <int BM, int BN, int BK, int WM, int WN, int KG = 1>
mfx_status launch_cfg(Handle* h, const F32Params& p, int grid, hipStream_t s) {
  // 128-bit loads only when both operands allow them.
  const bool vec = p.vec_a && p.vec_b;
  return timed_launch(h, "gemm_f32_mfma", s, [&] {
    if(vec) launch_layout<BM, BN, BK, WM, WN, KG, true>(p, grid, s);
    else launch_layout<BM, BN, BK, WM, WN, KG, false>(p, grid, s);
  });
}

template <int BM, int BN, int BK, int WM, int WN, int KG, bool VEC>
void launch_split_layout(const F32SplitParams& p, int grid, hipStream_t s) {
  constexpr int NT = F32Cfg<BM, BN, BK, WM, WN, KG>::NT;
  const bool ak = p.op.a_kmajor, bk = p.op.b_kmajor;
  if(ak && !bk) gemm_f32_splitk_kernel<BM, BN, BK, WM, WN, KG, true, false, VEC><<<grid, NT, 0, s>>>(p);
  else if(ak && bk) gemm_f32_splitk_kernel<BM, BN, BK, WM, WN, KG, true, true, VEC><<<grid, NT, 0, s>>>(p);
  else if(!ak && !bk) gemm_f32_splitk_kernel<BM, BN, BK, WM, WN, KG, false, false, VEC><<<grid, NT, 0, s>>>(p);
  else gemm_f32_splitk_kernel<BM, BN, BK, WM, WN, KG, false, true, VEC><<<grid, NT, 0, s>>>(p);
}

}  // namespace

// k slices per tile for a full single-product FP32 problem, 1 for none: only
// for the small-tile configurations (their partial tiles are 16-64 KiB);
// aims at sixteen waves per CU (four workgroups of four waves, one of
// sixteen), each slice at least four k stages long (two for the sixteen-wave
// tile, whose stages are deeper), and at most 32 MiB of partials.  Measured
// on MI355X (FP32, square and long-K shapes from 512 to 2048): the sixteen-
// wave tile unsplit at 1024^3, four slices at 512^3 and 512 x 512 x 8192.
int f32_splits(int config, int64_t tiles, int64_t K, int cus) {
  if(config < 0 || config >= kF32ConfigCount) return 1;
  const F32Config& c = kF32Configs[config];
  if(c.bm * c.bn > 128 * 128 || tiles <= 0) return 1;
  const int64_t per_cu = std::max(1, 16 / c.waves);
  int64_t s = std::min<int64_t>(8, per_cu * cus / tiles);
  s = std::min<int64_t>(s, K / ((c.waves >= 16 ? 2 : 4) * c.bk));
  while(s > 1 && tiles * s * c.bm * c.bn * 4 > ((int64_t)32 << 20)) --s;
  return (int)std::max<int64_t>(1, s);
}

size_t f32_splitk_workspace(int config, int64_t tiles, int splits) {
  const F32Config& c = kF32Configs[config];
  return (size_t)tiles * splits * c.bm * c.bn * sizeof(float);
}

// The split-K launches (slice kernel, then reduction) on a caller-provided
// partial buffer of f32_splitk_workspace() bytes.
mfx_status launch_f32_splitk(Handle* h, const DenseOp& op, int config, int splits, float* partial, hipStream_t s) {
  const F32Config& c = kF32Configs[config];
  F32SplitParams p;
  p.op = op;
  p.tiles_m = (int)ceil_div(op.M, c.bm);
  p.tiles_n = (int)ceil_div(op.N, c.bn);
  p.group = 8;
  p.kslice = (int)round_up(ceil_div(op.K, splits), c.bk);
  p.splits = (int)ceil_div(op.K, p.kslice);  // every slice non-empty
  p.partial = partial;
  const bool vec = f32_vec_ok(op.A, op.lda, op.a_kmajor, op.M, op.K) && f32_vec_ok(op.B, op.ldb, op.b_kmajor, op.N, op.K);
  const int grid = p.tiles_m * p.tiles_n * p.splits;
  const int rgrid = (int)ceil_div((int64_t)op.M * ceil_div(op.N, 4), 256);
  MFX_CHECK(timed_launch(h, "gemm_f32_splitk", s, [&] {
    switch(config) {
      case 1:
        if(vec) launch_split_layout<128, 128, 16, 64, 64, 1, true>(p, grid, s);
        else launch_split_layout<128, 128, 16, 64, 64, 1, false>(p, grid, s);
        break;
      case 2:
        if(vec) launch_split_layout<64, 64, 32, 32, 32, 1, true>(p, grid, s);
        else launch_split_layout<64, 64, 32, 32, 32, 1, false>(p, grid, s);
        break;
      case 6:
        if(vec) launch_split_layout<128, 128, 16, 64, 32, 1, true>(p, grid, s);
        else launch_split_layout<128, 128, 16, 64, 32, 1, false>(p, grid, s);
        break;
      case 7:
        if(vec) launch_split_layout<64, 64, 64, 32, 32, 4, true>(p, grid, s);
        else launch_split_layout<64, 64, 64, 32, 32, 4, false>(p, grid, s);
        break;
    }
  }));
  return timed_launch(h, "splitk_reduce", s, [&] {
    if(c.bm == 64) splitk_reduce_kernel<64, 64><<<rgrid, 256, 0, s>>>(p);
    else splitk_reduce_kernel<128, 128><<<rgrid, 256, 0, s>>>(p);
  });
}

int f32_config_count() { return kF32ConfigCount; }
bool f32_config_info(int c, int* bm, int* bn, int* bk, int* waves) {
  if(c < 0 || c >= kF32ConfigCount) return false;
  *bm = kF32Configs[c].bm;
  *bn = kF32Configs[c].bn;
  *bk = kF32Configs[c].bk;
  *waves = kF32Configs[c].waves;
  return true;
}

static mfx_status launch_f32_config(Handle* h, const F32Params& p, int config, int grid, hipStream_t s) {
  switch(config) {
    case 0: return launch_cfg<256, 256, 16, 128, 128>(h, p, grid, s);
    case 1: return launch_cfg<128, 128, 16, 64, 64>(h, p, grid, s);
    case 2: return launch_cfg<64, 64, 32, 32, 32>(h, p, grid, s);
    case 3: return launch_cfg<256, 128, 16, 128, 64>(h, p, grid, s);
    case 4: return launch_cfg<256, 256, 16, 128, 64>(h, p, grid, s);
    case 5: return launch_cfg<256, 128, 16, 64, 64>(h, p, grid, s);
    case 6: return launch_cfg<128, 128, 16, 64, 32>(h, p, grid, s);
    case 7: return launch_cfg<64, 64, 64, 32, 32, 4>(h, p, grid, s);
  }
  return invalid("unknown FP32 configuration");
}

// k slices for this call: the handle's override, else the automatic choice;
// 1 for triangular schedules and two-product (syr2k) contractions.
static int f32_call_splits(const Handle* h, const DenseOp& op, int config, int64_t tiles) {
  if(op.tiles != TILES_FULL || op.nprod != 1 || h->forced_splits == 1) return 1;
  const int cus = h->cu_count > 0 ? h->cu_count : 256;
  if(h->forced_splits < 0) return f32_splits(config, tiles, op.K, cus);
  const F32Config& c = kF32Configs[config];
  if(c.bm * c.bn > 128 * 128) return 1;
  int64_t s = std::min<int64_t>(h->forced_splits, op.K / c.bk);  // each slice at least one k stage
  while(s > 1 && tiles * s * c.bm * c.bn * 4 > ((int64_t)32 << 20)) --s;
  return (int)std::max<int64_t>(1, s);
}

mfx_status launch_dense_f32(Handle* h, const DenseOp& op, int config, hipStream_t s) {
  return launch_dense_f32_ws(h, op, config, nullptr, 0, s);
}

mfx_status launch_dense_f32_ws(Handle* h, const DenseOp& op, int config, void* ws, size_t ws_bytes, hipStream_t s) {
  if(config < 0 || config >= kF32ConfigCount) return invalid("FP32 tile configuration out of range");
  F32Params p;
  p.op = op;
  const F32Config& c = kF32Configs[config];
  p.tiles_m = (int)ceil_div(op.M, c.bm);
  p.tiles_n = (int)ceil_div(op.N, c.bn);
  p.group = 8;
  p.vec_a = f32_vec_ok(op.A, op.lda, op.a_kmajor, op.M, op.K);
  p.vec_b = f32_vec_ok(op.B, op.ldb, op.b_kmajor, op.N, op.K);
  if(op.nprod == 2) {
    // second product swaps the operands: both must allow the same load width
    const bool v = p.vec_a && p.vec_b;
    p.vec_a = p.vec_b = v;
  }
  int grid;
  TriTail tail;
  if(op.tiles == TILES_FULL) {
    grid = p.tiles_m * p.tiles_n;
  } else {
    if(p.tiles_m != p.tiles_n || c.bm != c.bn) return invalid("triangular schedule needs square tiles");
    grid = p.tiles_m * (p.tiles_m + 1) / 2;
    tail = tri_tail(MFX_PREC_FP32, config, p.tiles_m, h->cu_count > 0 ? h->cu_count : 256);
  }
  if(grid == 0) return MFX_SUCCESS;
  if(tail.small >= 0) {
    // the full rounds on this tile, the last round's big tiles re-tiled
    MFX_CHECK(launch_f32_config(h, p, config, tail.head, s));
    F32Params q = p;
    const int sb = kF32Configs[tail.small].bm;
    q.tiles_m = q.tiles_n = (int)ceil_div(op.M, sb);
    q.tail_first = tail.head;
    q.tail_ratio = tail.ratio;
    q.big_tiles = p.tiles_m;
    return launch_f32_config(h, q, tail.small, (grid - tail.head) * tail.ratio * tail.ratio, s);
  }
  const int splits = f32_call_splits(h, op, config, grid);
  if(splits > 1) {
    const size_t need = f32_splitk_workspace(config, grid, splits);
    if(ws) {
      if(need > ws_bytes) return unsupported("split-K partial tiles exceed the given buffer");
    } else {
      MFX_CHECK(workspace(h, need, &ws));
    }
    return launch_f32_splitk(h, op, config, splits, static_cast<float*>(ws), s);
  }
  return launch_f32_config(h, p, config, grid, s);
}

}  // namespace mfx
