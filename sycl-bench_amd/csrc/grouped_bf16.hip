// Persistent grouped kernels for split / single BF16: every item is one
// output tile, computed by bf16_tile_direct straight from the FP32 operands
// (the hi/lo split happens while staging, so there is no packing pass and no
// packed copy to wait for).  As for FP32, the 256x256 configuration is not
// compiled here: next to the persistent loop its 256 accumulators and the
// staging registers do not fit without spills, so 256x128 is the large tile.
#include "grouped.h"

namespace mfx {
namespace {

template <int NHL, int C, bool VEC>
__device__ __forceinline__ void run_tile(const DenseOp& op, const GProblem& P, int tm, int tn, u32x4* lds) {
  constexpr int BM = Geo<C>::BM, BN = Geo<C>::BN, WM = Geo<C>::WM, WN = Geo<C>::WN;
  if(P.a_kmajor && !P.b_kmajor) bf16_tile_direct<BM, BN, WM, WN, NHL, true, false, VEC>(op, tm, tn, lds);
  else if(P.a_kmajor) bf16_tile_direct<BM, BN, WM, WN, NHL, true, true, VEC>(op, tm, tn, lds);
  else if(!P.b_kmajor) bf16_tile_direct<BM, BN, WM, WN, NHL, false, false, VEC>(op, tm, tn, lds);
  else bf16_tile_direct<BM, BN, WM, WN, NHL, false, true, VEC>(op, tm, tn, lds);
}

template <int NHL, int C>
constexpr int bf16_smem_chunks() {
  return BfCfg<Geo<C>::BM, Geo<C>::BN, Geo<C>::WM, Geo<C>::WN, NHL>::SMEM_CHUNKS;
}

// VEC: every problem of the launch allows 128-bit operand loads.
template <int NHL, int BIG, int SMALL, bool VEC>
__global__ __launch_bounds__(256) void grouped_bf16x_kernel(const GroupedArgs g) {
  constexpr int S0 = bf16_smem_chunks<NHL, BIG>(), S1 = bf16_smem_chunks<NHL, SMALL>();
  __shared__ u32x4 lds[S0 > S1 ? S0 : S1];
  __shared__ int slot;
  for(;;) {
    const int it = take_item(g.head, &slot);
    if(it >= g.nitems) break;
    const GItem item = g.items[it];
    const GProblem& P = g.problems[item.problem];
    if(!wait_producers(P, g)) continue;
    DenseOp op{};
    op.M = P.M;
    op.N = P.N;
    op.K = P.K;
    op.A = P.A;
    op.lda = P.lda;
    op.a_kmajor = P.a_kmajor;
    op.B = P.B;
    op.ldb = P.ldb;
    op.b_kmajor = P.b_kmajor;
    op.nprod = 1;
    op.tiles = TILES_FULL;
    op.ep = Epilogue{P.C, P.ldc, P.alpha, P.beta, P.M, P.N, TILES_FULL, 0};
    if constexpr(BIG == SMALL) {
      run_tile<NHL, BIG, VEC>(op, P, item.x, item.y, lds);
    } else {
      if(P.config == BIG) run_tile<NHL, BIG, VEC>(op, P, item.x, item.y, lds);
      else run_tile<NHL, SMALL, VEC>(op, P, item.x, item.y, lds);
    }
    signal(&g.done[item.problem]);
  }
}

template <int NHL, bool VEC>
bool launch_pair(int big, int small, const GroupedArgs& g, int grid, hipStream_t s) {
  switch(big * 4 + small) {
    case 3 * 4 + 3: grouped_bf16x_kernel<NHL, 3, 3, VEC><<<grid, 256, 0, s>>>(g); return true;
    case 3 * 4 + 1: grouped_bf16x_kernel<NHL, 3, 1, VEC><<<grid, 256, 0, s>>>(g); return true;
    case 3 * 4 + 2: grouped_bf16x_kernel<NHL, 3, 2, VEC><<<grid, 256, 0, s>>>(g); return true;
    case 1 * 4 + 1: grouped_bf16x_kernel<NHL, 1, 1, VEC><<<grid, 256, 0, s>>>(g); return true;
    case 1 * 4 + 2: grouped_bf16x_kernel<NHL, 1, 2, VEC><<<grid, 256, 0, s>>>(g); return true;
    case 2 * 4 + 2: grouped_bf16x_kernel<NHL, 2, 2, VEC><<<grid, 256, 0, s>>>(g); return true;
  }
  return false;
}

}  // namespace

mfx_status launch_grouped_bf16(Handle* h, int nhl, int big, int small, bool vec, const GroupedArgs& g, int grid,
    hipStream_t s) {
  bool known = false;
  const mfx_status st = timed_launch(h, nhl == 2 ? "grouped_split_bf16_mfma" : "grouped_bf16_mfma", s, [&] {
    if(nhl == 2) known = vec ? launch_pair<2, true>(big, small, g, grid, s) : launch_pair<2, false>(big, small, g, grid, s);
    else known = vec ? launch_pair<1, true>(big, small, g, grid, s) : launch_pair<1, false>(big, small, g, grid, s);
  });
  if(st == MFX_SUCCESS && !known) return invalid("no grouped BF16 kernel for this tile pair");
  return st;
}

}  // namespace mfx
