// Persistent grouped launch: many uneven contractions, and chains of them
// (2mm, 3mm, matmulchain), in one kernel launch.
//
// The host turns the problem list into a work queue of items in topological
// order:
//   PACK_A / PACK_B (split / single BF16 only): convert one tile-row panel of
//       an operand into fragment order (tile_bf16.h) in the workspace;
//   TILE: one output tile of one problem.
// Level by level (a problem's level is 1 + the highest level of its
// producers): the level's pack items, then its tile items, largest K first.
// A fixed grid of one workgroup per CU pulls items from an atomic head
// counter.  Before working on an item it waits, with a bounded spin, for
//   - PACK: every producer of the problem to have finished all its tiles;
//   - TILE: its A and B panels to be packed (BF16 modes), or its producers to
//     be finished (FP32, which reads the operands directly).
// Completion is published by fence + barrier + atomic increment; waits are
// acquire loads at agent scope, so producer results written on another XCD
// are visible.  Deadlock freedom: a workgroup only waits for items that
// precede its own in the queue, and each of those was taken by a workgroup
// that is running and waits only for items before it.  A wait that exceeds
// its spin limit sets an error flag and skips the item; every workgroup still
// drains the queue and exits.
#include <algorithm>
#include <cstdlib>
#include <numeric>

#include "tile_bf16.h"
#include "tile_f32.h"

namespace mfx {

int choose_grouped_config(Handle* h, mfx_precision p, int64_t m, int64_t n, int64_t k);

namespace {

constexpr int MAX_DEPS = 4;
enum ItemType : int { ITEM_PACK_A = 0, ITEM_PACK_B = 1, ITEM_TILE = 2 };

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
  int config;
  int tiles_m, tiles_n;
  int kbs;                 // BF16: padded K / 16
  u32x4* packed_a;         // BF16: fragment-order operands
  u32x4* packed_b;
  int pack_a_flags;        // offset of this problem's A-panel flags in the flag array
  int pack_b_flags;
  int ndeps;
  int deps[MAX_DEPS];      // producer problem indices
};

struct GItem {
  int type, problem, x, y;  // PACK: x = panel; TILE: (x, y) = (tm, tn)
};

struct GroupedArgs {
  const GProblem* problems;
  const GItem* items;
  int nitems;
  int* head;        // next item to take
  int* error;       // set to 1 when a wait times out
  int* done;        // per problem: finished tiles
  int* pack_flags;  // per packed panel: 1 when packed
  int spin_limit;
};

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

__device__ __forceinline__ int take_item(int* head, int* slot) {
  __syncthreads();  // everyone is done with the previous item (and its LDS)
  if(threadIdx.x == 0) *slot = atomicAdd(head, 1);
  __syncthreads();
  return *slot;
}

__device__ __forceinline__ int tile_bm(int config) { return config == 1 ? 128 : config == 2 ? 64 : 256; }
__device__ __forceinline__ int tile_bn(int config) { return config == 0 ? 256 : config == 2 ? 64 : 128; }

template <int NHL>
__device__ void pack_panel(const GProblem& P, bool is_a, int panel) {
  const int rows_per_panel = is_a ? tile_bm(P.config) : tile_bn(P.config);
  PackSrc s;
  s.src = is_a ? P.A : P.B;
  s.ld = is_a ? P.lda : P.ldb;
  s.rows = is_a ? P.M : P.N;
  s.K = P.K;
  s.kmajor = is_a ? P.a_kmajor : P.b_kmajor;
  s.vec = is_a ? P.vec_a : P.vec_b;
  u32x4* dst = is_a ? P.packed_a : P.packed_b;
  const int blk0 = panel * rows_per_panel / 32;
  const int chunks = rows_per_panel / 32 * P.kbs * 64;
  for(int c = threadIdx.x; c < chunks; c += 256) {
    const int lane = c & 63, t = c >> 6;
    pack_chunk<NHL>(s, blk0 + t / P.kbs, t % P.kbs, lane, P.kbs, dst);
  }
}

template <int NHL>
__global__ __launch_bounds__(256) void grouped_bf16x_kernel(const GroupedArgs g) {
  __shared__ u32x4 lds[BfCfg<256, 256, 128, 128, NHL>::SMEM_CHUNKS];
  __shared__ int slot;
  for(;;) {
    const int it = take_item(g.head, &slot);
    if(it >= g.nitems) break;
    const GItem item = g.items[it];
    const GProblem& P = g.problems[item.problem];
    if(item.type != ITEM_TILE) {
      const bool is_a = item.type == ITEM_PACK_A;
      if(!wait_producers(P, g)) continue;
      pack_panel<NHL>(P, is_a, item.x);
      signal(&g.pack_flags[(is_a ? P.pack_a_flags : P.pack_b_flags) + item.x]);
      continue;
    }
    const bool ready = wait_at_least(&g.pack_flags[P.pack_a_flags + item.x], 1, g) &&
                       wait_at_least(&g.pack_flags[P.pack_b_flags + item.y], 1, g);
    if(!ready) continue;
    BfTileArgs t;
    t.PA = P.packed_a;
    t.PB = P.packed_b;
    t.kbs = P.kbs;
    t.nprod = 1;
    t.ep = Epilogue{P.C, P.ldc, P.alpha, P.beta, P.M, P.N, TILES_FULL, 0};
    switch(P.config) {
      case 0: bf16_tile<256, 256, 128, 128, NHL>(t, item.x, item.y, lds); break;
      case 1: bf16_tile<128, 128, 64, 64, NHL>(t, item.x, item.y, lds); break;
      case 2: bf16_tile<64, 64, 32, 32, NHL>(t, item.x, item.y, lds); break;
      default: bf16_tile<256, 128, 128, 64, NHL>(t, item.x, item.y, lds); break;
    }
    signal(&g.done[item.problem]);
  }
}

template <int BM, int BN, int BK, int WM, int WN>
__device__ __forceinline__ void f32_tile_any(const DenseOp& op, const GProblem& P, int tm, int tn, float* smem) {
  if(P.a_kmajor && !P.b_kmajor) f32_tile<BM, BN, BK, WM, WN, true, false>(op, P.vec_a, P.vec_b, tm, tn, smem);
  else if(P.a_kmajor) f32_tile<BM, BN, BK, WM, WN, true, true>(op, P.vec_a, P.vec_b, tm, tn, smem);
  else if(!P.b_kmajor) f32_tile<BM, BN, BK, WM, WN, false, false>(op, P.vec_a, P.vec_b, tm, tn, smem);
  else f32_tile<BM, BN, BK, WM, WN, false, true>(op, P.vec_a, P.vec_b, tm, tn, smem);
}

__global__ __launch_bounds__(256) void grouped_f32_kernel(const GroupedArgs g) {
  __shared__ float smem[F32Cfg<256, 256, 16, 128, 128>::SMEM_FLOATS];
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
    switch(P.config) {
      case 0: f32_tile_any<256, 256, 16, 128, 128>(op, P, item.x, item.y, smem); break;
      case 1: f32_tile_any<128, 128, 16, 64, 64>(op, P, item.x, item.y, smem); break;
      case 2: f32_tile_any<64, 64, 32, 32, 32>(op, P, item.x, item.y, smem); break;
      default: f32_tile_any<256, 128, 16, 128, 64>(op, P, item.x, item.y, smem); break;
    }
    signal(&g.done[item.problem]);
  }
}

// Clears the queue counters and the error flag before each launch (a kernel,
// so the reset can be part of a captured graph).
__global__ void zero_ints_kernel(int* p, int n, int* err) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
  if(blockIdx.x == 0 && threadIdx.x == 0) *err = 0;
}

int host_bm(int config) { return config == 1 ? 128 : config == 2 ? 64 : 256; }
int host_bn(int config) { return config == 0 ? 256 : config == 2 ? 64 : 128; }

// Bytes of the packed A and B copies of one problem (each rounded to 256).
std::pair<size_t, size_t> packed_sizes(const GProblem& P, int nhl) {
  const size_t ba = (size_t)P.tiles_m * host_bm(P.config) * P.kbs * 16 * nhl * 2;
  const size_t bb = (size_t)P.tiles_n * host_bn(P.config) * P.kbs * 16 * nhl * 2;
  return {(size_t)round_up(ba, 256), (size_t)round_up(bb, 256)};
}

}  // namespace

// A grouped launch prepared on the host: tables, workspace layout and the
// arguments of the persistent kernel.
struct GroupedPlan {
  mfx_precision prec;
  std::vector<GProblem> problems;
  std::vector<GItem> items;
  size_t table_bytes = 0, counter_ints = 0, packed_bytes = 0;
  size_t off_problems = 0, off_items = 0, off_counters = 0, off_packed = 0;
  size_t total_bytes = 0;
  int n_flags = 0;
  double flops = 0.0;
};

static mfx_status validate(const mfx_gemm_problem* problems, int count, const mfx_dep* deps, int ndeps) {
  if(count < 0 || ndeps < 0) return invalid("negative problem or dependency count");
  if(count > 0 && !problems) return invalid("null problem list");
  if(ndeps > 0 && !deps) return invalid("null dependency list");
  for(int i = 0; i < count; ++i) {
    const auto& q = problems[i];
    if(q.m < 0 || q.n < 0 || q.k < 0 || q.m > (1 << 30) || q.n > (1 << 30) || q.k > (1 << 30))
      return invalid("problem " + std::to_string(i) + ": bad dimensions");
    if((q.transA != 0 && q.transA != 1) || (q.transB != 0 && q.transB != 1))
      return invalid("problem " + std::to_string(i) + ": transA / transB must be 0 or 1");
    const int64_t a_rows = q.transA ? q.k : q.m, a_cols = q.transA ? q.m : q.k;
    const int64_t b_rows = q.transB ? q.n : q.k, b_cols = q.transB ? q.k : q.n;
    if(q.m > 0 && q.n > 0) {
      if(!q.C || q.ldc < std::max<int64_t>(1, q.n)) return invalid("problem " + std::to_string(i) + ": bad C");
      if(q.k > 0 && (!q.A || !q.B)) return invalid("problem " + std::to_string(i) + ": null operand");
      if(q.lda < std::max<int64_t>(1, a_cols) || q.ldb < std::max<int64_t>(1, b_cols))
        return invalid("problem " + std::to_string(i) + ": leading dimension smaller than row length");
    }
    (void)a_rows;
    (void)b_rows;
  }
  for(int d = 0; d < ndeps; ++d) {
    if(deps[d].consumer < 0 || deps[d].consumer >= count || deps[d].producer < 0 || deps[d].producer >= count)
      return invalid("dependency " + std::to_string(d) + " names a problem out of range");
    if(deps[d].producer >= deps[d].consumer)
      return invalid("dependency " + std::to_string(d) + ": producers must come before their consumers");
  }
  return MFX_SUCCESS;
}

mfx_status plan_grouped(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps, GroupedPlan& pl) {
  if(prec != MFX_PREC_FP32 && prec != MFX_PREC_SPLIT_BF16 && prec != MFX_PREC_BF16)
    return unsupported("grouped launch supports FP32, SPLIT_BF16 and BF16");
  MFX_CHECK(validate(problems, count, deps, ndeps));
  pl.prec = prec;
  const bool packed = prec != MFX_PREC_FP32;
  const int nhl = prec == MFX_PREC_SPLIT_BF16 ? 2 : 1;
  pl.problems.resize(count);
  std::vector<int> level(count, 0);
  for(int i = 0; i < count; ++i) pl.problems[i].ndeps = 0;
  for(int d = 0; d < ndeps; ++d) {
    GProblem& c = pl.problems[deps[d].consumer];
    bool dup = false;
    for(int e = 0; e < c.ndeps; ++e) dup |= c.deps[e] == deps[d].producer;
    if(dup) continue;
    if(c.ndeps == MAX_DEPS) return unsupported("a problem may depend on at most 4 producers");
    c.deps[c.ndeps++] = deps[d].producer;
  }
  for(int i = 0; i < count; ++i)
    for(int e = 0; e < pl.problems[i].ndeps; ++e) level[i] = std::max(level[i], level[pl.problems[i].deps[e]] + 1);

  size_t packed_bytes = 0;
  int flags = 0;
  for(int i = 0; i < count; ++i) {
    const auto& q = problems[i];
    GProblem& P = pl.problems[i];
    P.M = (int)q.m;
    P.N = (int)q.n;
    P.K = (int)q.k;
    P.alpha = q.alpha;
    P.beta = q.beta;
    P.A = q.A;
    P.lda = q.lda;
    P.B = q.B;
    P.ldb = q.ldb;
    P.C = q.C;
    P.ldc = q.ldc;
    P.a_kmajor = q.transA == 0;
    P.b_kmajor = q.transB == 1;
    P.config = choose_grouped_config(h, prec, q.m, q.n, q.k);
    const int bm = host_bm(P.config), bn = host_bn(P.config);
    P.tiles_m = (int)ceil_div(q.m, bm);
    P.tiles_n = (int)ceil_div(q.n, bn);
    if(q.m == 0 || q.n == 0) P.tiles_m = P.tiles_n = 0;
    P.vec_a = f32_vec_ok(q.A, q.lda, P.a_kmajor, P.M, P.K);
    P.vec_b = f32_vec_ok(q.B, q.ldb, P.b_kmajor, P.N, P.K);
    if(packed) {
      // pack_chunk only takes 128-bit loads along k
      P.vec_a = P.a_kmajor && P.vec_a;
      P.vec_b = P.b_kmajor && P.vec_b;
      P.kbs = (int)(round_up(q.k, 32) / 16);
      packed_bytes += packed_sizes(P, nhl).first + packed_sizes(P, nhl).second;
      P.packed_a = P.packed_b = nullptr;  // placed in bind_grouped
      P.pack_a_flags = flags;
      flags += P.tiles_m;
      P.pack_b_flags = flags;
      flags += P.tiles_n;
    } else {
      P.kbs = 0;
      P.packed_a = P.packed_b = nullptr;
      P.pack_a_flags = P.pack_b_flags = 0;
    }
    pl.flops += 2.0 * q.m * q.n * q.k;
  }
  pl.n_flags = flags;

  // Items level by level: packs, then tiles (largest K first, a problem's
  // tiles row by row).
  const int max_level = count ? *std::max_element(level.begin(), level.end()) : 0;
  for(int lv = 0; lv <= max_level; ++lv) {
    std::vector<int> at;
    for(int i = 0; i < count; ++i)
      if(level[i] == lv && pl.problems[i].tiles_m * pl.problems[i].tiles_n > 0) at.push_back(i);
    if(packed)
      for(int i : at) {
        for(int p = 0; p < pl.problems[i].tiles_m; ++p) pl.items.push_back({ITEM_PACK_A, i, p, 0});
        for(int p = 0; p < pl.problems[i].tiles_n; ++p) pl.items.push_back({ITEM_PACK_B, i, p, 0});
      }
    std::stable_sort(at.begin(), at.end(), [&](int a, int b) { return pl.problems[a].K > pl.problems[b].K; });
    for(int i : at)
      for(int tm = 0; tm < pl.problems[i].tiles_m; ++tm)
        for(int tn = 0; tn < pl.problems[i].tiles_n; ++tn) pl.items.push_back({ITEM_TILE, i, tm, tn});
  }

  pl.off_problems = 0;
  pl.off_items = round_up(sizeof(GProblem) * std::max(1, count), 256);
  pl.off_counters = pl.off_items + round_up(sizeof(GItem) * std::max<size_t>(1, pl.items.size()), 256);
  pl.counter_ints = 2 + (size_t)count + (size_t)flags;
  pl.off_packed = pl.off_counters + round_up(pl.counter_ints * sizeof(int), 256);
  pl.packed_bytes = packed_bytes;
  pl.total_bytes = pl.off_packed + packed_bytes;
  return MFX_SUCCESS;
}

// Places the packed operands in the device buffer `ws`, uploads the tables
// (stream-ordered, from pinned staging memory owned by the handle) and fills
// the kernel arguments.
static mfx_status bind_grouped(Handle* h, GroupedPlan& pl, void* ws, GroupedArgs& g, hipStream_t s) {
  char* base = static_cast<char*>(ws);
  if(pl.prec != MFX_PREC_FP32) {
    const int nhl = pl.prec == MFX_PREC_SPLIT_BF16 ? 2 : 1;
    size_t off = pl.off_packed;
    for(auto& P : pl.problems) {
      const auto sz = packed_sizes(P, nhl);
      P.packed_a = reinterpret_cast<u32x4*>(base + off);
      off += sz.first;
      P.packed_b = reinterpret_cast<u32x4*>(base + off);
      off += sz.second;
    }
  }
  // The previous upload from the staging buffer must have completed.
  if(h->staging_ready) MFX_HIP(hipEventSynchronize(h->staging_ready));
  if(h->staging_bytes < pl.off_counters) {
    if(h->staging) MFX_HIP(hipHostFree(h->staging));
    h->staging = nullptr;
    h->staging_bytes = 0;
    MFX_HIP(hipHostMalloc(&h->staging, pl.off_counters, hipHostMallocDefault));
    h->staging_bytes = pl.off_counters;
  }
  if(!h->staging_ready) MFX_HIP(hipEventCreateWithFlags(&h->staging_ready, hipEventDisableTiming));
  char* st = static_cast<char*>(h->staging);
  memcpy(st + pl.off_problems, pl.problems.data(), sizeof(GProblem) * pl.problems.size());
  memcpy(st + pl.off_items, pl.items.data(), sizeof(GItem) * pl.items.size());
  MFX_HIP(hipMemcpyAsync(base, st, pl.off_counters, hipMemcpyHostToDevice, s));
  MFX_HIP(hipEventRecord(h->staging_ready, s));
  int* counters = reinterpret_cast<int*>(base + pl.off_counters);
  g.problems = reinterpret_cast<const GProblem*>(base + pl.off_problems);
  g.items = reinterpret_cast<const GItem*>(base + pl.off_items);
  g.nitems = (int)pl.items.size();
  g.head = counters;
  g.error = counters + 1;
  g.done = counters + 2;
  g.pack_flags = counters + 2 + pl.problems.size();
  g.spin_limit = h->spin_limit;
  return MFX_SUCCESS;
}

// Enqueues the counter reset and the persistent kernel.
static mfx_status launch_grouped_kernels(Handle* h, const GroupedPlan& pl, const GroupedArgs& g, hipStream_t s) {
  const int n = (int)pl.counter_ints;
  MFX_CHECK(timed_launch(h, "grouped_reset", s, [&] {
    zero_ints_kernel<<<(int)std::min<int64_t>(ceil_div(n, 256), 1024), 256, 0, s>>>(g.head, n, g.error);
  }));
  if(pl.items.empty()) return MFX_SUCCESS;
  const int grid = h->cu_count > 0 ? h->cu_count : 256;
  switch(pl.prec) {
    case MFX_PREC_FP32:
      return timed_launch(h, "grouped_f32_mfma", s, [&] { grouped_f32_kernel<<<grid, 256, 0, s>>>(g); });
    case MFX_PREC_SPLIT_BF16:
      return timed_launch(h, "grouped_split_bf16_mfma", s, [&] { grouped_bf16x_kernel<2><<<grid, 256, 0, s>>>(g); });
    default:
      return timed_launch(h, "grouped_bf16_mfma", s, [&] { grouped_bf16x_kernel<1><<<grid, 256, 0, s>>>(g); });
  }
}

mfx_status run_grouped(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps, hipStream_t s) {
  GroupedPlan pl;
  MFX_CHECK(plan_grouped(h, prec, problems, count, deps, ndeps, pl));
  void* ws = nullptr;
  MFX_CHECK(workspace(h, pl.total_bytes, &ws));
  GroupedArgs g;
  MFX_CHECK(bind_grouped(h, pl, ws, g, s));
  // The flag outlives the workspace (which later calls may regrow).
  if(!h->error_flag) MFX_HIP(hipMalloc(&h->error_flag, 256));
  g.error = h->error_flag;
  return launch_grouped_kernels(h, pl, g, s);
}

// Balanced-tree product of `count` n x n matrices as grouped problems.
mfx_status chain_problems(int count, const float* const* mats, int64_t n, float* const* scratch, float* out,
    std::vector<mfx_gemm_problem>& probs, std::vector<mfx_dep>& deps) {
  if(count < 2) return invalid("a chain needs at least two matrices");
  if(!mats || !out || (count > 2 && !scratch)) return invalid("null matrix list");
  struct Node {
    const float* p;
    int producer;  // problem index, -1 for an input
  };
  std::vector<Node> level;
  for(int i = 0; i < count; ++i) {
    if(!mats[i]) return invalid("null matrix in chain");
    level.push_back({mats[i], -1});
  }
  int next_scratch = 0;
  while(level.size() > 1) {
    std::vector<Node> up;
    for(size_t i = 0; i + 1 < level.size(); i += 2) {
      const bool last = level.size() == 2;
      float* dst = last ? out : scratch[next_scratch++];
      if(!dst) return invalid("null scratch buffer");
      const int idx = (int)probs.size();
      probs.push_back({n, n, n, 1.0f, 0.0f, level[i].p, n, 0, level[i + 1].p, n, 0, dst, n});
      if(level[i].producer >= 0) deps.push_back({idx, level[i].producer});
      if(level[i + 1].producer >= 0) deps.push_back({idx, level[i + 1].producer});
      up.push_back({dst, idx});
    }
    if(level.size() % 2) up.push_back(level.back());
    level.swap(up);
  }
  return MFX_SUCCESS;
}

}  // namespace mfx

using namespace mfx;

struct mfx_graph_s {
  int device = 0;
  void* mem = nullptr;  // tables, counters and packed panels of this graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int* error = nullptr;
};

extern "C" {

mfx_status mfx_gemm_grouped(mfx_handle handle, mfx_precision prec, const mfx_gemm_problem* problems, int32_t count,
    const mfx_dep* deps, int32_t ndeps, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  return run_grouped(h, prec, problems, count, deps, ndeps, static_cast<hipStream_t>(stream));
}

mfx_status mfx_chain(mfx_handle handle, mfx_precision prec, int32_t count, const float* const* mats, int64_t n,
    float* const* scratch, float* out, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  if(n < 0) return invalid("negative dimension");
  std::vector<mfx_gemm_problem> probs;
  std::vector<mfx_dep> deps;
  MFX_CHECK(chain_problems(count, mats, n, scratch, out, probs, deps));
  return run_grouped(h, prec, probs.data(), (int)probs.size(), deps.data(), (int)deps.size(),
      static_cast<hipStream_t>(stream));
}

mfx_status mfx_grouped_status(mfx_handle handle, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  MFX_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  if(!h->error_flag) return MFX_SUCCESS;
  int err = 0;
  MFX_HIP(hipMemcpy(&err, h->error_flag, sizeof(int), hipMemcpyDeviceToHost));
  if(err) {
    set_last_error("an in-kernel dependency wait of the last grouped launch exceeded its spin limit");
    return MFX_ERROR_TIMEOUT;
  }
  return MFX_SUCCESS;
}

mfx_status mfx_grouped_graph_create(mfx_handle handle, mfx_precision prec, const mfx_gemm_problem* problems,
    int32_t count, const mfx_dep* deps, int32_t ndeps, mfx_stream stream, mfx_graph* graph) {
  if(!handle || !graph) return invalid("null handle or graph pointer");
  *graph = nullptr;
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  GroupedPlan pl;
  MFX_CHECK(plan_grouped(h, prec, problems, count, deps, ndeps, pl));
  auto* gr = new(std::nothrow) mfx_graph_s;
  if(!gr) return MFX_ERROR_ALLOC;
  gr->device = h->device;
  // The graph owns its tables, counters and packed panels, so later calls on
  // the handle (which may regrow the handle's workspace) cannot move them.
  hipError_t ae = hipMalloc(&gr->mem, std::max<size_t>(pl.total_bytes, 256));
  if(ae != hipSuccess) {
    delete gr;
    return hip_fail(ae, "hipMalloc(grouped graph)");
  }
  GroupedArgs g;
  mfx_status bs = bind_grouped(h, pl, gr->mem, g, s);
  hipError_t se = hipStreamSynchronize(s);  // tables resident before capture
  if(bs != MFX_SUCCESS || se != hipSuccess) {
    (void)hipFree(gr->mem);
    delete gr;
    return bs != MFX_SUCCESS ? bs : hip_fail(se, "hipStreamSynchronize");
  }
  gr->error = g.error;
  // Capture on a private stream: the graph holds the counter reset (a
  // kernel, not a memset node) and the persistent kernel.
  hipStream_t cs = nullptr;
  hipError_t ce = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if(ce != hipSuccess) {
    (void)hipFree(gr->mem);
    delete gr;
    return hip_fail(ce, "hipStreamCreateWithFlags");
  }
  const bool timing = h->timing;
  h->timing = false;
  h->capturing = true;
  mfx_status st = MFX_SUCCESS;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if(e == hipSuccess) {
    st = launch_grouped_kernels(h, pl, g, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &gr->graph);
    if(st == MFX_SUCCESS && e2 != hipSuccess) st = hip_fail(e2, "hipStreamEndCapture");
  } else {
    st = hip_fail(e, "hipStreamBeginCapture");
  }
  h->capturing = false;
  h->timing = timing;
  if(st == MFX_SUCCESS) {
    e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
    if(e != hipSuccess) st = hip_fail(e, "hipGraphInstantiate");
  }
  (void)hipStreamDestroy(cs);
  if(st != MFX_SUCCESS) {
    if(gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if(gr->graph) (void)hipGraphDestroy(gr->graph);
    (void)hipFree(gr->mem);
    delete gr;
    return st;
  }
  *graph = gr;
  return MFX_SUCCESS;
}

mfx_status mfx_graph_launch(mfx_graph graph, mfx_stream stream) {
  if(!graph) return invalid("null graph");
  MFX_HIP(hipSetDevice(graph->device));
  MFX_HIP(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
  return MFX_SUCCESS;
}

mfx_status mfx_graph_status(mfx_graph graph, mfx_stream stream) {
  if(!graph) return invalid("null graph");
  MFX_HIP(hipSetDevice(graph->device));
  MFX_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  int err = 0;
  MFX_HIP(hipMemcpy(&err, graph->error, sizeof(int), hipMemcpyDeviceToHost));
  if(err) {
    set_last_error("an in-kernel dependency wait of the last graph replay exceeded its spin limit");
    return MFX_ERROR_TIMEOUT;
  }
  return MFX_SUCCESS;
}

mfx_status mfx_graph_destroy(mfx_graph graph) {
  if(!graph) return MFX_SUCCESS;
  hipError_t e = hipSetDevice(graph->device);
  if(e == hipSuccess) e = hipDeviceSynchronize();  // no replay may still use the memory
  if(graph->exec && e == hipSuccess) e = hipGraphExecDestroy(graph->exec);
  if(graph->graph && e == hipSuccess) e = hipGraphDestroy(graph->graph);
  if(graph->mem && e == hipSuccess) e = hipFree(graph->mem);
  delete graph;
  return e == hipSuccess ? MFX_SUCCESS : hip_fail(e, "mfx_graph_destroy");
}

}  // extern "C"
