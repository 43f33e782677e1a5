// Persistent grouped launch: tables shared by the host planner (grouped.hip)
// and the kernels (grouped_bf16.hip, grouped_f32.hip).
//
// The host turns the problem list into a work queue of items in topological
// order, level by level (a problem's level is 1 + the highest level of its
// producers):
//   PACK_A / PACK_B (split / single BF16 only): split one 32-row block x
//       PACK_KBLOCKS k blocks of an operand into the fragment order of
//       tile_bf16.h, in the workspace;
//   TILE: one output tile of one problem.
// Each level lists its pack items, then its tiles, largest K first.  A fixed
// grid of one workgroup per CU pulls items from an atomic head counter.
// Before working on an item it waits, with a bounded spin, for
//   - PACK: every producer of the problem to have finished all its tiles;
//   - TILE (BF16): every pack item of its A and B panels (a counter per
//     panel); TILE (FP32, which reads the operands directly): the producers.
// Pack items are small (about 1-2 us each) so a level's packing spreads over
// all workgroups at HBM speed instead of one workgroup per panel.
// Completion is published by fence + barrier + atomic increment; a wait is
// polled by one thread with relaxed loads and ends in an agent-scope
// acquire, so producer results written on another XCD are visible.
// Deadlock freedom: a workgroup only waits for items that precede its own in
// the queue, and each of those was taken by a workgroup that is running and
// waits only for items before it.  A wait that exceeds
// its spin limit sets an error flag and skips the item; every other wait
// gives up at its next poll once the flag is set, so every workgroup drains
// the queue and exits promptly.
//
// Tile shapes: a launch runs one tile policy, a large and a small tile
// configuration; each problem takes the large tile unless that pads its
// output area by more than the policy's bound.  A kernel holds the bodies of
// at most two configurations: with more, the register allocator spills into
// scratch memory inside the k loop.
#pragma once

#include "tile_bf16.h"
#include "tile_f32.h"

namespace mfx {

constexpr int GROUPED_MAX_DEPS = 4;

constexpr int PACK_KBLOCKS = 32;  // 16-wide k blocks per pack item (512 k)

// ITEM_PREPACK (split / single BF16, one launch per segment): pack blocks
// [x, x + y) of the segment's pre-pack job list -- operands of problems that
// the NEXT segment's launch contracts, so no item of this launch waits for it
// and the kernel boundary orders it before its readers.
enum GroupedItemType : int { ITEM_PACK_A = 0, ITEM_PACK_B = 1, ITEM_TILE = 2, ITEM_PREPACK = 3 };

struct GProblem {
  int M, N, K;
  float alpha, beta;
  const float* A;
  int64_t lda;
  const float* B;
  int64_t ldb;
  float* C;
  int64_t ldc;
  int a_kmajor, b_kmajor;
  int vec_a, vec_b;
  int config;  // tile configuration of this problem
  int tiles_m, tiles_n;
  int kbs;                    // BF16: padded K / 16
  int pack_slices;            // BF16: pack items per 32-row block (ceil(kbs / PACK_KBLOCKS))
  u32x4* packed_a;            // BF16: fragment-order operands
  u32x4* packed_b;
  int pack_a_ctr, pack_b_ctr; // BF16: first panel counter of A / B
  int prepacked;              // BF16: operands packed by the pre-pass kernel
  int direct;                 // BF16: tiles read the FP32 operands directly (bf16_tile_direct), no packing
  int consumed;               // another problem waits for this one's tiles
  int ndeps;
  int deps[GROUPED_MAX_DEPS];  // producer problem indices
};

// One operand packed by the pre-pass kernel (problems without producers):
// chunks [first_block*256, (first_block+blocks)*256) of the flattened job list.
struct PackJob {
  PackSrc src;
  u32x4* dst;
  int kbs;
  int blocks;       // 256-chunk blocks of this job (the last one partly used)
  int64_t chunks;   // rows_pad / 32 * kbs * 64
  int first_block;
};

struct GItem {
  int type, problem, x, y;  // TILE: (x, y) = (tm, tn); PACK: x = 32-row block, y = k slice
};

// Work queues per launch: one per XCD.  Workgroups are dispatched
// round-robin over the 8 XCDs, so workgroup b drains queue b % 8 first (the
// planner puts all tiles of a problem in one queue, so the tiles that share
// an operand panel run on one XCD and meet that panel in its L2), then the
// other queues in turn.  A launch whose items wait on each other in-kernel
// puts every item in one queue, which keeps the queue order.
constexpr int GROUPED_QUEUES = 8;

struct GroupedArgs {
  const GProblem* problems;
  const GItem* items;
  int nitems;
  int* head;        // GROUPED_QUEUES counters: items taken from each queue
  const int* qbeg;  // queue q holds items [qbeg[q], qbeg[q + 1])
  int* error;       // set to 1 when a wait times out
  int* done;        // per problem: finished tiles
  int* pack_ctr;    // per packed panel: finished pack items
  int spin_limit;
  const PackJob* prejobs;  // ITEM_PREPACK: the segment's pre-pack jobs (blocks numbered from 0)
  int nprejobs;
};

// Tile geometry of configuration C (the same ids as the single-problem
// kernels): BM x BN per workgroup, WM x WN per wave (four waves for 0-3,
// eight, two per SIMD, for 4-6); BK is the FP32 k-step.
template <int C> struct Geo;
template <> struct Geo<0> { static constexpr int BM = 256, BN = 256, WM = 128, WN = 128, BK = 16; };
template <> struct Geo<1> { static constexpr int BM = 128, BN = 128, WM = 64, WN = 64, BK = 16; };
template <> struct Geo<2> { static constexpr int BM = 64, BN = 64, WM = 32, WN = 32, BK = 32; };
template <> struct Geo<3> { static constexpr int BM = 256, BN = 128, WM = 128, WN = 64, BK = 16; };
template <> struct Geo<4> { static constexpr int BM = 256, BN = 256, WM = 128, WN = 64, BK = 16; };
template <> struct Geo<5> { static constexpr int BM = 256, BN = 128, WM = 64, WN = 64, BK = 16; };
template <> struct Geo<6> { static constexpr int BM = 128, BN = 128, WM = 64, WN = 32, BK = 16; };
// threads of a workgroup running configuration C
template <int C>
constexpr int geo_threads() {
  return (Geo<C>::BM / Geo<C>::WM) * (Geo<C>::BN / Geo<C>::WN) * 64;
}

__host__ __device__ inline int cfg_bm(int config) { return config == 1 || config == 6 ? 128 : config == 2 ? 64 : 256; }
__host__ __device__ inline int cfg_bn(int config) {
  return config == 0 || config == 4 ? 256 : config == 2 ? 64 : 128;
}
// four waves per workgroup, or eight for configurations 4-6
__host__ __device__ inline int cfg_threads(int config) { return config >= 4 ? 512 : 256; }

// Cross-XCD ordering costs cache maintenance on gfx950: an agent-scope
// release writes back the XCD's L2 (buffer_wbl2) and an agent-scope acquire
// invalidates it (buffer_inv).  So waits poll with relaxed atomic loads
// (performed at the coherent level, no invalidation) and acquire once when
// the condition holds, and a signal writes back once per workgroup.

// Bounded poll for *ctr >= target, by one thread.  It gives up as soon as
// any workgroup has set the error flag (a wait elsewhere timed out, so the
// item this one waits for may never be published), or after spin_limit
// polls, when it sets the flag itself.  So one expired wait ends every other
// wait of the launch within a poll instead of each spinning its own limit.
__device__ __forceinline__ bool poll_at_least(int* ctr, int target, const GroupedArgs& g) {
  int spins = 0;
  while(__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if(__hip_atomic_load(g.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if(++spins > g.spin_limit) {
      __hip_atomic_store(g.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(8);
  }
  return true;
}

// Polls every producer of P in turn, stopping at the first that fails.
__device__ __forceinline__ bool poll_producers(const GProblem& P, const GroupedArgs& g) {
  for(int d = 0; d < P.ndeps; ++d) {
    const GProblem& Q = g.problems[P.deps[d]];
    if(!poll_at_least(&g.done[P.deps[d]], Q.tiles_m * Q.tiles_n, g)) return false;
  }
  return true;
}

// Workgroup-uniform wait: thread 0 runs `poll` (the other waves do not add
// polling traffic), the answer is shared through `flag` in LDS, so the whole
// workgroup either runs the item or skips it -- never half of it, which
// would leave the tile's barriers unmatched.  On success thread 0's
// agent-scope acquire makes the producers' results visible to it, and the
// barrier orders every other thread's later loads after that acquire.
template <class Poll>
__device__ __forceinline__ bool wg_wait(int* flag, Poll&& poll) {
  if(threadIdx.x == 0) {
    const bool ok = poll();
    if(ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *flag = ok ? 1 : 0;
  }
  __syncthreads();
  return *flag != 0;
}

// Publishes the workgroup's global writes, then bumps *ctr once: every thread
// waits for its own stores to reach L2, the barrier orders them before
// thread 0, whose agent-scope release writes the L2 back before the add.
__device__ __forceinline__ void signal(int* ctr) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if(threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Next item index for the whole workgroup, -1 when every queue is empty.
// `tried` (LDS, zero at kernel start) counts the queues this workgroup has
// found empty: it drains its own XCD's queue, then the next, and so on.
// The value is the same in every lane; reading it through readfirstlane lets
// the compiler keep the item and its problem in scalar registers, leaving
// the vector registers to the tile.
__device__ __forceinline__ int take_item(const GroupedArgs& g, int* slot, int* tried) {
  __syncthreads();  // everyone is done with the previous item (and its LDS)
  if(threadIdx.x == 0) {
    int it = -1;
    while(*tried < GROUPED_QUEUES) {
      const int q = (int)((blockIdx.x + *tried) % GROUPED_QUEUES);
      const int len = g.qbeg[q + 1] - g.qbeg[q];
      if(len > 0) {
        const int k = atomicAdd(&g.head[q], 1);
        if(k < len) {
          it = g.qbeg[q] + k;
          break;
        }
      }
      ++*tried;
    }
    *slot = it;
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*slot);
}

// A grouped launch prepared on the host: tables, workspace layout and the
// arguments of the persistent kernel.
struct GroupedPlan {
  mfx_precision prec;
  std::vector<GProblem> problems;
  std::vector<GItem> items;
  std::vector<int> qbeg;  // per segment GROUPED_QUEUES + 1 queue bounds (relative to the segment's items)
  size_t counter_ints = 0;
  size_t off_problems = 0, off_items = 0, off_qbeg = 0, off_jobs = 0, off_counters = 0, off_packed = 0;
  // Pack jobs: the pre-pass kernels' (jobs[0, npre)) then the pre-pack
  // items' (jobs[npre, ...)); a job's first_block counts from the start of
  // its segment's range.  job_op: (problem, 1 = A / 0 = B) of each job.
  std::vector<PackJob> jobs;
  std::vector<std::pair<int, int>> job_op;
  int npre = 0;
  // Launch segments: one for the whole queue, or one per dependency level
  // (per_level), or pipelined segments.  Segment l covers items
  // [seg_items[l], seg_items[l+1]), pre-pass jobs [seg_jobs[l],
  // seg_jobs[l+1]) in seg_blocks[l] blocks, and its pre-pack items pack jobs
  // [seg_prejobs[l], seg_prejobs[l+1]) (indices into jobs).
  bool per_level = false, pipelined = false;
  std::vector<int> seg_items, seg_jobs, seg_blocks, seg_prejobs;
  std::vector<int> seg_big, seg_small;  // tile pair of each segment's kernel
  const PackJob* dev_jobs = nullptr;          // the job table in device memory
  size_t total_bytes = 0;
  int policy = 0, big = 0, small = 0;  // tile policy and its configurations
  bool all_vec = true;                 // every problem allows 128-bit operand loads
  double flops = 0.0;
};

// Plans a grouped launch (host only: no device memory is touched).
mfx_status plan_grouped(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps, GroupedPlan& pl);

// Plans and binds a grouped launch once, launches it once untimed, then
// returns in *ms the average time of `reps` back-to-back launches of the same
// bound tables (what a captured graph replays: no host planning between
// launches, so the GPU never idles between them).
mfx_status time_grouped(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps, int reps, hipStream_t s, double* ms);

// Pre-pass: packs every job's operand (one chunk per thread).
mfx_status launch_pack_jobs(Handle* h, int nhl, const PackJob* jobs, int njobs, int blocks, hipStream_t s);

// Kernel launchers (one instantiation per pair of configurations).
// vec: every problem allows 128-bit operand loads.
mfx_status launch_grouped_bf16(Handle* h, int nhl, int big, int small, bool vec, const GroupedArgs& g, int grid,
    hipStream_t s);
mfx_status launch_grouped_f32(Handle* h, int big, int small, bool vec, const GroupedArgs& g, int grid,
    hipStream_t s);

}  // namespace mfx
