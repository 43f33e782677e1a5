// Persistent grouped launch: tables shared by the host planner (grouped.hip)
// and the kernels (grouped_bf16.hip, grouped_f32.hip).
//
// The host turns the problem list into a work queue of items, one per output
// tile, in topological order: level by level (a problem's level is 1 + the
// highest level of its producers), largest K first.  A fixed grid of one
// workgroup per CU pulls items from an atomic head counter.  Before working
// on a tile it waits, with a bounded spin, for every producer of its problem
// to have finished all its tiles.  Tiles read the FP32 operands directly; the
// BF16 modes split them while staging (bf16_tile_direct).
// Completion is published by fence + barrier + atomic increment; waits are
// acquire loads at agent scope, so producer results written on another XCD
// are visible.  Deadlock freedom: a workgroup only waits for items that
// precede its own in the queue, and each of those was taken by a workgroup
// that is running and waits only for items before it.  A wait that exceeds
// its spin limit sets an error flag and skips the item; every workgroup still
// drains the queue and exits.
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
  int ndeps;
  int deps[GROUPED_MAX_DEPS];  // producer problem indices
};

struct GItem {
  int problem, x, y, pad;  // output tile (x, y) = (tm, tn) of `problem`
};

struct GroupedArgs {
  const GProblem* problems;
  const GItem* items;
  int nitems;
  int* head;        // next item to take
  int* error;       // set to 1 when a wait times out
  int* done;        // per problem: finished tiles
  int spin_limit;
};

// Tile geometry of configuration C (the same ids as the single-problem
// kernels): BM x BN per workgroup of four waves, WM x WN per wave; BK is the
// FP32 k-step.
template <int C> struct Geo;
template <> struct Geo<0> { static constexpr int BM = 256, BN = 256, WM = 128, WN = 128, BK = 16; };
template <> struct Geo<1> { static constexpr int BM = 128, BN = 128, WM = 64, WN = 64, BK = 16; };
template <> struct Geo<2> { static constexpr int BM = 64, BN = 64, WM = 32, WN = 32, BK = 32; };
template <> struct Geo<3> { static constexpr int BM = 256, BN = 128, WM = 128, WN = 64, BK = 16; };

inline int host_bm(int config) { return config == 1 ? 128 : config == 2 ? 64 : 256; }
inline int host_bn(int config) { return config == 0 ? 256 : config == 2 ? 64 : 128; }

// Bounded wait for *ctr >= target (all threads of the workgroup).
__device__ __forceinline__ bool wait_at_least(int* ctr, int target, const GroupedArgs& g) {
  int spins = 0;
  while(__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if(++spins > g.spin_limit) {
      __hip_atomic_store(g.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(8);
  }
  return true;
}

__device__ __forceinline__ bool wait_producers(const GProblem& P, const GroupedArgs& g) {
  bool ok = true;
  for(int d = 0; d < P.ndeps; ++d) {
    const GProblem& Q = g.problems[P.deps[d]];
    ok = wait_at_least(&g.done[P.deps[d]], Q.tiles_m * Q.tiles_n, g) && ok;
  }
  return ok;
}

// Publishes the workgroup's global writes, then bumps *ctr once.
__device__ __forceinline__ void signal(int* ctr) {
  __threadfence();
  __syncthreads();
  if(threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Next item index for the whole workgroup.  The value is the same in every
// lane; reading it through readfirstlane lets the compiler keep the item and
// its problem in scalar registers, leaving the vector registers to the tile.
__device__ __forceinline__ int take_item(int* head, int* slot) {
  __syncthreads();  // everyone is done with the previous item (and its LDS)
  if(threadIdx.x == 0) *slot = atomicAdd(head, 1);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*slot);
}

// Kernel launchers (one instantiation per pair of configurations).
// vec: every problem allows 128-bit operand loads.
mfx_status launch_grouped_bf16(Handle* h, int nhl, int big, int small, bool vec, const GroupedArgs& g, int grid,
    hipStream_t s);
mfx_status launch_grouped_f32(Handle* h, int big, int small, bool vec, const GroupedArgs& g, int grid,
    hipStream_t s);

}  // namespace mfx
