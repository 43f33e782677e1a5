// Persistent grouped kernels for FP32: tile items run f32_tile directly on
// the operands (no packing).  The 256x256 configuration is not compiled here:
// next to the persistent loop its 256 accumulators and prefetch registers do
// not fit without spills, so FP32 policies use 256x128 as the large tile.
#include "grouped.h"

namespace mfx {
namespace {

template <int C, bool VEC>
__device__ __forceinline__ void run_tile(const DenseOp& op, const GProblem& P, int tm, int tn, float* smem) {
  constexpr int BM = Geo<C>::BM, BN = Geo<C>::BN, BK = Geo<C>::BK, WM = Geo<C>::WM, WN = Geo<C>::WN;
  if(P.a_kmajor && !P.b_kmajor) f32_tile<BM, BN, BK, WM, WN, true, false, VEC>(op, tm, tn, smem);
  else if(P.a_kmajor) f32_tile<BM, BN, BK, WM, WN, true, true, VEC>(op, tm, tn, smem);
  else if(!P.b_kmajor) f32_tile<BM, BN, BK, WM, WN, false, false, VEC>(op, tm, tn, smem);
  else f32_tile<BM, BN, BK, WM, WN, false, true, VEC>(op, tm, tn, smem);
}

template <int C>
constexpr int f32_smem_floats() {
  return F32Cfg<Geo<C>::BM, Geo<C>::BN, Geo<C>::BK, Geo<C>::WM, Geo<C>::WN>::SMEM_FLOATS;
}

// VEC: every problem of the launch allows 128-bit operand loads.
template <int BIG, int SMALL, bool VEC>
__global__ __launch_bounds__(geo_threads<BIG>()) void grouped_f32_kernel(const GroupedArgs g) {
  static_assert(geo_threads<SMALL>() == geo_threads<BIG>(), "both tiles of a policy use the same workgroup");
  constexpr int S0 = f32_smem_floats<BIG>(), S1 = f32_smem_floats<SMALL>();
  __shared__ float smem[S0 > S1 ? S0 : S1];
  __shared__ int slot, ready, tried;
  if(threadIdx.x == 0) tried = 0;
  for(;;) {
    const int it = take_item(g, &slot, &tried);
    if(it < 0) break;
    const GItem item = g.items[it];
    const GProblem& P = g.problems[item.problem];
    if(P.ndeps > 0 && !wg_wait(&ready, [&] { return poll_producers(P, g); })) continue;
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
      run_tile<BIG, VEC>(op, P, item.x, item.y, smem);
    } else {
      if(P.config == BIG) run_tile<BIG, VEC>(op, P, item.x, item.y, smem);
      else run_tile<SMALL, VEC>(op, P, item.x, item.y, smem);
    }
    if(P.consumed) signal(&g.done[item.problem]);
  }
}

template <int BIG, int SMALL, bool VEC>
void launch_one(const GroupedArgs& g, int grid, hipStream_t s) {
  grouped_f32_kernel<BIG, SMALL, VEC><<<grid, geo_threads<BIG>(), 0, s>>>(g);
}

template <bool VEC>
bool launch_pair(int big, int small, const GroupedArgs& g, int grid, hipStream_t s) {
  switch(big * 8 + small) {
    case 3 * 8 + 3: launch_one<3, 3, VEC>(g, grid, s); return true;
    case 3 * 8 + 1: launch_one<3, 1, VEC>(g, grid, s); return true;
    case 3 * 8 + 2: launch_one<3, 2, VEC>(g, grid, s); return true;
    case 1 * 8 + 1: launch_one<1, 1, VEC>(g, grid, s); return true;
    case 1 * 8 + 2: launch_one<1, 2, VEC>(g, grid, s); return true;
    case 2 * 8 + 2: launch_one<2, 2, VEC>(g, grid, s); return true;
    case 4 * 8 + 4: launch_one<4, 4, VEC>(g, grid, s); return true;
    case 4 * 8 + 5: launch_one<4, 5, VEC>(g, grid, s); return true;
    case 4 * 8 + 6: launch_one<4, 6, VEC>(g, grid, s); return true;
    case 5 * 8 + 5: launch_one<5, 5, VEC>(g, grid, s); return true;
    case 5 * 8 + 6: launch_one<5, 6, VEC>(g, grid, s); return true;
    case 6 * 8 + 6: launch_one<6, 6, VEC>(g, grid, s); return true;
  }
  return false;
}

}  // namespace

mfx_status launch_grouped_f32(Handle* h, int big, int small, bool vec, const GroupedArgs& g, int grid,
    hipStream_t s) {
  bool known = false;
  const mfx_status st = timed_launch(h, "grouped_f32_mfma", s, [&] {
    known = vec ? launch_pair<true>(big, small, g, grid, s) : launch_pair<false>(big, small, g, grid, s);
  });
  if(st == MFX_SUCCESS && !known) return invalid("no grouped FP32 kernel for this tile pair");
  return st;
}

}  // namespace mfx
