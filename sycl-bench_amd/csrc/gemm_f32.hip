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
};

template <int BM, int BN, int BK, int WM, int WN, bool AK, bool BKM, bool VEC>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_f32_kernel(const F32Params p) {
  __shared__ float smem[F32Cfg<BM, BN, BK, WM, WN>::SMEM_FLOATS];
  int tm, tn;
  const int pid = xcd_remap(blockIdx.x, gridDim.x);
  if(p.op.tiles == TILES_FULL) {
    grouped_tile(pid, p.tiles_m, p.tiles_n, p.group, tm, tn);
  } else {
    int ti, tj;
    lower_tile(pid, ti, tj);
    if(p.op.tiles == TILES_LOWER) { tm = ti; tn = tj; } else { tm = tj; tn = ti; }
  }
  f32_tile<BM, BN, BK, WM, WN, AK, BKM, VEC>(p.op, tm, tn, smem);
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
};
constexpr int kF32ConfigCount = sizeof(kF32Configs) / sizeof(kF32Configs[0]);

template <int BM, int BN, int BK, int WM, int WN, bool VEC>
void launch_layout(const F32Params& p, int grid, hipStream_t s) {
  constexpr int NT = F32Cfg<BM, BN, BK, WM, WN>::NT;
  const bool ak = p.op.a_kmajor, bk = p.op.b_kmajor;
  if(ak && !bk) gemm_f32_kernel<BM, BN, BK, WM, WN, true, false, VEC><<<grid, NT, 0, s>>>(p);
  else if(ak && bk) gemm_f32_kernel<BM, BN, BK, WM, WN, true, true, VEC><<<grid, NT, 0, s>>>(p);
  else if(!ak && !bk) gemm_f32_kernel<BM, BN, BK, WM, WN, false, false, VEC><<<grid, NT, 0, s>>>(p);
  else gemm_f32_kernel<BM, BN, BK, WM, WN, false, true, VEC><<<grid, NT, 0, s>>>(p);
}

template <int BM, int BN, int BK, int WM, int WN>
mfx_status launch_cfg(Handle* h, const F32Params& p, int grid, hipStream_t s) {
  // 128-bit loads only when both operands allow them.
  const bool vec = p.vec_a && p.vec_b;
  return timed_launch(h, "gemm_f32_mfma", s, [&] {
    if(vec) launch_layout<BM, BN, BK, WM, WN, true>(p, grid, s);
    else launch_layout<BM, BN, BK, WM, WN, false>(p, grid, s);
  });
}

}  // namespace

int f32_config_count() { return kF32ConfigCount; }
bool f32_config_info(int c, int* bm, int* bn, int* bk, int* waves) {
  if(c < 0 || c >= kF32ConfigCount) return false;
  *bm = kF32Configs[c].bm;
  *bn = kF32Configs[c].bn;
  *bk = kF32Configs[c].bk;
  *waves = kF32Configs[c].waves;
  return true;
}

mfx_status launch_dense_f32(Handle* h, const DenseOp& op, int config, hipStream_t s) {
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
  if(op.tiles == TILES_FULL) {
    grid = p.tiles_m * p.tiles_n;
  } else {
    if(p.tiles_m != p.tiles_n || c.bm != c.bn) return invalid("triangular schedule needs square tiles");
    grid = p.tiles_m * (p.tiles_m + 1) / 2;
  }
  if(grid == 0) return MFX_SUCCESS;
  switch(config) {
    case 0: return launch_cfg<256, 256, 16, 128, 128>(h, p, grid, s);
    case 1: return launch_cfg<128, 128, 16, 64, 64>(h, p, grid, s);
    case 2: return launch_cfg<64, 64, 32, 32, 32>(h, p, grid, s);
    case 3: return launch_cfg<256, 128, 16, 128, 64>(h, p, grid, s);
    case 4: return launch_cfg<256, 256, 16, 128, 64>(h, p, grid, s);
    case 5: return launch_cfg<256, 128, 16, 64, 64>(h, p, grid, s);
    case 6: return launch_cfg<128, 128, 16, 64, 32>(h, p, grid, s);
  }
  return invalid("unknown FP32 configuration");
}

}  // namespace mfx
