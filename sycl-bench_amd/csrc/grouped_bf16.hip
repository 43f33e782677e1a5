// Grouped kernels for split / single BF16.  A pre-pass kernel packs the
// operands of every problem without producers (the kernel boundary orders it
// before the tiles, with no in-kernel synchronisation).  In the persistent
// kernel, pack items split one 32-row block x PACK_KBLOCKS k blocks of an
// operand of a problem with producers once those are done; tile items run
// bf16_tile (LDS stages filled by global-to-LDS DMA) on the packed operands,
// for consumer problems once every pack item of their panels is published.
#include "grouped.h"

namespace mfx {
namespace {

// One pack item: all loads of a thread are issued before any of its stores
// (the stores could alias the loads as far as the compiler knows, so it
// would not reorder them itself).
template <int NHL, int NT>
__device__ void pack_item(const GProblem& P, bool is_a, int blk, int slice) {
  PackSrc s;
  s.src = is_a ? P.A : P.B;
  s.ld = is_a ? P.lda : P.ldb;
  s.rows = is_a ? P.M : P.N;
  s.K = P.K;
  s.kmajor = is_a ? P.a_kmajor : P.b_kmajor;
  s.vec = is_a ? P.vec_a : P.vec_b;
  u32x4* dst = is_a ? P.packed_a : P.packed_b;
  const int kb0 = slice * PACK_KBLOCKS;
  const int nkb = min(PACK_KBLOCKS, P.kbs - kb0);
  constexpr int PER = PACK_KBLOCKS * 64 / NT;
  float v[PER][8];
#pragma unroll
  for(int i = 0; i < PER; ++i) {
    const int c = threadIdx.x + i * NT;
    if((c >> 6) < nkb) load_chunk8(s, blk, kb0 + (c >> 6), c & 63, v[i]);
  }
#pragma unroll
  for(int i = 0; i < PER; ++i) {
    const int c = threadIdx.x + i * NT;
    if((c >> 6) < nkb) store_chunk<NHL>(dst, blk, kb0 + (c >> 6), c & 63, P.kbs, v[i]);
  }
}

// Pre-pass: block b packs chunks of the job whose block range holds b (jobs
// start on 256-chunk blocks, so each wave's 64 chunks are one (blk, kb)
// pair, as load_chunk8_staged needs; a non-k-major source walks the pairs
// kb-major, neighbouring waves on neighbouring row segments).
template <int NHL>
__global__ __launch_bounds__(256) void pack_jobs_kernel(const PackJob* jobs, int njobs) {
  __shared__ float tiles[4][STAGED_TILE_FLOATS];
  int lo = 0, hi = njobs - 1;
  const int b = blockIdx.x;
  while(lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if(jobs[mid].first_block <= b) lo = mid;
    else hi = mid - 1;
  }
  const PackJob& j = jobs[lo];
  const int64_t c = (int64_t)(b - j.first_block) * 256 + threadIdx.x;
  const bool active = c < j.chunks;
  const int lane = (int)(c & 63);
  const int64_t t = c >> 6;
  if(j.src.kmajor) {
    if(active) pack_chunk<NHL>(j.src, (int)(t / j.kbs), (int)(t % j.kbs), lane, j.kbs, j.dst);
    return;
  }
  const int nblk = (int)(j.chunks / ((int64_t)j.kbs * 64));
  const int blk = (int)(t % nblk), kb = (int)(t / nblk);
  float v[8];
  load_chunk8_staged(j.src, blk, kb, lane, active, v, tiles[threadIdx.x >> 6]);
  if(active) store_chunk<NHL>(j.dst, blk, kb, lane, j.kbs, v);
}

template <int NHL, int C>
__device__ __forceinline__ void run_tile(const BfTileArgs& t, int tm, int tn, u32x4* lds) {
  bf16_tile<Geo<C>::BM, Geo<C>::BN, Geo<C>::WM, Geo<C>::WN, NHL>(t, tm, tn, lds);
}

template <int NHL, int C>
constexpr int bf16_smem_chunks() {
  return BfCfg<Geo<C>::BM, Geo<C>::BN, Geo<C>::WM, Geo<C>::WN, NHL>::SMEM_CHUNKS;
}

template <int NHL, int BIG, int SMALL>
__global__ __launch_bounds__(geo_threads<BIG>()) void grouped_bf16x_kernel(const GroupedArgs g) {
  constexpr int NT = geo_threads<BIG>();
  static_assert(geo_threads<SMALL>() == NT, "both tiles of a policy use the same workgroup");
  constexpr int S0 = bf16_smem_chunks<NHL, BIG>(), S1 = bf16_smem_chunks<NHL, SMALL>();
  __shared__ u32x4 lds[S0 > S1 ? S0 : S1];
  __shared__ int slot, ready;
  for(;;) {
    const int it = take_item(g.head, &slot);
    if(it >= g.nitems) break;
    const GItem item = g.items[it];
    const GProblem& P = g.problems[item.problem];
    const int bm = cfg_bm(P.config), bn = cfg_bn(P.config);
    if(item.type != ITEM_TILE) {
      const bool is_a = item.type == ITEM_PACK_A;
      if(P.ndeps > 0 && !wg_wait(&ready, [&] { return poll_producers(P, g); })) continue;
      pack_item<NHL, NT>(P, is_a, item.x, item.y);
      const int panel = item.x * 32 / (is_a ? bm : bn);
      signal(&g.pack_ctr[(is_a ? P.pack_a_ctr : P.pack_b_ctr) + panel]);
      continue;
    }
    const int nblk_a = (P.M + 31) / 32, nblk_b = (P.N + 31) / 32;
    if(!P.prepacked && P.pack_slices > 0) {
      // pack items of a panel: its 32-row blocks that meet the matrix, each in pack_slices k slices
      const int need_a = (min((item.x + 1) * (bm / 32), nblk_a) - item.x * (bm / 32)) * P.pack_slices;
      const int need_b = (min((item.y + 1) * (bn / 32), nblk_b) - item.y * (bn / 32)) * P.pack_slices;
      if(!wg_wait(&ready, [&] {
           return poll_at_least(&g.pack_ctr[P.pack_a_ctr + item.x], need_a, g) &&
                  poll_at_least(&g.pack_ctr[P.pack_b_ctr + item.y], need_b, g);
         }))
        continue;
    } else if(!P.prepacked && P.ndeps > 0) {
      // K == 0: no pack items, but the tile still rewrites C (beta * C), so
      // it waits for the producers like any other consumer
      if(!wg_wait(&ready, [&] { return poll_producers(P, g); })) continue;
    }
    BfTileArgs t;
    t.PA = P.packed_a;
    t.PB = P.packed_b;
    t.kbs = P.kbs;
    t.nblk_a = nblk_a;
    t.nblk_b = nblk_b;
    t.nprod = 1;
    t.ep = Epilogue{P.C, P.ldc, P.alpha, P.beta, P.M, P.N, TILES_FULL, 0};
    if constexpr(BIG == SMALL) {
      run_tile<NHL, BIG>(t, item.x, item.y, lds);
    } else {
      if(P.config == BIG) run_tile<NHL, BIG>(t, item.x, item.y, lds);
      else run_tile<NHL, SMALL>(t, item.x, item.y, lds);
    }
    if(P.consumed) signal(&g.done[item.problem]);
  }
}

template <int NHL, int BIG, int SMALL>
void launch_one(const GroupedArgs& g, int grid, hipStream_t s) {
  grouped_bf16x_kernel<NHL, BIG, SMALL><<<grid, geo_threads<BIG>(), 0, s>>>(g);
}

template <int NHL>
bool launch_pair(int big, int small, const GroupedArgs& g, int grid, hipStream_t s) {
  switch(big * 8 + small) {
    case 0 * 8 + 0: launch_one<NHL, 0, 0>(g, grid, s); return true;
    case 0 * 8 + 1: launch_one<NHL, 0, 1>(g, grid, s); return true;
    case 0 * 8 + 2: launch_one<NHL, 0, 2>(g, grid, s); return true;
    case 3 * 8 + 3: launch_one<NHL, 3, 3>(g, grid, s); return true;
    case 3 * 8 + 1: launch_one<NHL, 3, 1>(g, grid, s); return true;
    case 3 * 8 + 2: launch_one<NHL, 3, 2>(g, grid, s); return true;
    case 1 * 8 + 1: launch_one<NHL, 1, 1>(g, grid, s); return true;
    case 1 * 8 + 2: launch_one<NHL, 1, 2>(g, grid, s); return true;
    case 2 * 8 + 2: launch_one<NHL, 2, 2>(g, grid, s); return true;
    case 4 * 8 + 4: launch_one<NHL, 4, 4>(g, grid, s); return true;
    case 4 * 8 + 5: launch_one<NHL, 4, 5>(g, grid, s); return true;
    case 4 * 8 + 6: launch_one<NHL, 4, 6>(g, grid, s); return true;
    case 5 * 8 + 5: launch_one<NHL, 5, 5>(g, grid, s); return true;
    case 5 * 8 + 6: launch_one<NHL, 5, 6>(g, grid, s); return true;
    case 6 * 8 + 6: launch_one<NHL, 6, 6>(g, grid, s); return true;
  }
  return false;
}

}  // namespace

mfx_status launch_pack_jobs(Handle* h, int nhl, const PackJob* jobs, int njobs, int blocks, hipStream_t s) {
  if(njobs == 0 || blocks == 0) return MFX_SUCCESS;
  return timed_launch(h, nhl == 2 ? "grouped_pack_split_bf16" : "grouped_pack_bf16", s, [&] {
    if(nhl == 2) pack_jobs_kernel<2><<<blocks, 256, 0, s>>>(jobs, njobs);
    else pack_jobs_kernel<1><<<blocks, 256, 0, s>>>(jobs, njobs);
  });
}

mfx_status launch_grouped_bf16(Handle* h, int nhl, int big, int small, bool vec, const GroupedArgs& g, int grid,
    hipStream_t s) {
  (void)vec;  // pack items choose their load width per operand
  bool known = false;
  const mfx_status st = timed_launch(h, nhl == 2 ? "grouped_split_bf16_mfma" : "grouped_bf16_mfma", s, [&] {
    known = nhl == 2 ? launch_pair<2>(big, small, g, grid, s) : launch_pair<1>(big, small, g, grid, s);
  });
  if(st == MFX_SUCCESS && !known) return invalid("no grouped BF16 kernel for this tile pair");
  return st;
}

}  // namespace mfx
