// Host side of the persistent grouped launch (queue layout and protocol in
// grouped.h): planning, workspace binding, the tile policy, chains, the
// hipGraph form and the C entry points.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <queue>

#include "grouped.h"

namespace mfx {

double config_rate(mfx_precision p, int c);

std::string tune_key(mfx_precision p, int64_t m, int64_t n, int64_t k);
int tuned_lookup(const std::string& key);

namespace {

// Tile policies: (large, small, bound) -- a problem takes the large tile when
// that pads its output area by at most `bound` (0: always), else the small
// one.  The FP32 grouped kernels do not compile the 256x256 configuration (0)
// and run 256x128 (3) in its place.
struct GroupedPolicy {
  int big, small;
  double bound;
  // 0: one persistent kernel, in-kernel dependency waits; 1: one pre-pass +
  // persistent launch per dependency level, no in-kernel waits; 2:
  // pipelined segments -- like 1, with the independent problems spread over
  // at least three segments and each launch packing the next segment's
  // independent operands in pre-pack items beside its own tiles.
  int per_level;
  // BF16: problems whose operands allow it (K % 16 == 0, aligned rows) run
  // bf16_tile_direct on their FP32 operands instead of being packed.
  int direct;
};
constexpr GroupedPolicy kPolicies[] = {
    {0, 0, 0.0},   // 0
    {0, 1, 1.34},  // 1
    {0, 2, 1.34},  // 2
    {3, 3, 0.0},   // 3
    {3, 1, 1.34},  // 4
    {3, 1, 2.0},   // 5
    {3, 2, 1.34},  // 6
    {1, 1, 0.0},   // 7
    {1, 2, 1.34},  // 8
    {2, 2, 0.0},   // 9
    {4, 4, 0.0},   // 10: eight-wave tiles from here on
    {4, 6, 1.34},  // 11
    {4, 5, 1.34},  // 12
    {5, 5, 0.0},   // 13
    {5, 6, 1.34},  // 14
    {6, 6, 0.0},   // 15
    {4, 4, 0.0, 1},   // 16: one launch per dependency level from here on
    {4, 6, 1.34, 1},  // 17
    {4, 5, 1.34, 1},  // 18
    {3, 1, 1.34, 1},  // 19
    {6, 6, 0.0, 1},   // 20
    {4, 4, 0.0, 2},   // 21: pipelined segments from here on
    {4, 5, 1.34, 2},  // 22
    {4, 6, 1.34, 2},  // 23
    {3, 1, 1.34, 2},  // 24
    {4, 4, 0.0, 1, 1},   // 25: FP32 operands read directly (BF16) from here on
    {4, 5, 1.34, 1, 1},  // 26
    {4, 6, 1.34, 1, 1},  // 27
    {3, 1, 1.34, 1, 1},  // 28
    {4, 4, 0.0, 0, 1},   // 29: in-kernel waits
    {4, 5, 1.34, 0, 1},  // 30
    {4, 4, 0.0, 2, 1},   // 31: pipelined segments
    {4, 6, 1.34, 2, 1},  // 32
    {4, 5, 1.34, 2, 1},  // 33
    {4, 6, 0.0, 2, 1},   // 34: independent problems on the large tile, problems with
    {4, 5, 0.0, 2, 1},   // 35:   producers (a chain's later levels) on the small one
};
constexpr int kPolicyCount = sizeof(kPolicies) / sizeof(kPolicies[0]);

// Relative per-CU contraction rate of each tile configuration on large
// problems (larger tiles reuse each loaded panel more), used by the
// untuned policy estimate.
constexpr double kTileRate[7] = {1.0, 0.6, 0.2, 0.85, 1.15, 0.95, 0.65};

// The configuration a grouped launch of `prec` runs for configuration c.
int grouped_config(mfx_precision prec, int c) { return prec == MFX_PREC_FP32 && c == 0 ? 3 : c; }

int problem_config(const GroupedPolicy& pol, mfx_precision prec, int64_t m, int64_t n) {
  const int big = grouped_config(prec, pol.big), small = grouped_config(prec, pol.small);
  if(big == small || pol.bound <= 0.0) return big;
  const double padded = (double)round_up(m, cfg_bm(big)) * (double)round_up(n, cfg_bn(big));
  return padded <= pol.bound * (double)m * (double)n ? big : small;
}

// Topological level of every problem (producers come first, so one pass).
std::vector<int> problem_levels(int count, const mfx_dep* deps, int ndeps) {
  std::vector<int> level(count, 0);
  std::vector<std::vector<int>> prod(count);
  for(int d = 0; d < ndeps; ++d) prod[deps[d].consumer].push_back(deps[d].producer);
  for(int i = 0; i < count; ++i)
    for(int p : prod[i]) level[i] = std::max(level[i], level[p] + 1);
  return level;
}

// Tile configuration of every problem under a policy: the padding rule,
// then, on a dependency level whose tiles would not give every CU one, the
// small tile for all of the level's problems (a chain's later levels hold
// few problems, and a level cannot overlap the next).
std::vector<int> assign_configs(const GroupedPolicy& pol, mfx_precision prec, const mfx_gemm_problem* problems,
    int count, const std::vector<int>& level, int cus) {
  std::vector<int> cfg(count);
  for(int i = 0; i < count; ++i) cfg[i] = problem_config(pol, prec, problems[i].m, problems[i].n);
  const int small = grouped_config(prec, pol.small);
  if(grouped_config(prec, pol.big) == small) return cfg;
  const int levels = count ? *std::max_element(level.begin(), level.end()) + 1 : 0;
  std::vector<int64_t> tiles(levels, 0);
  for(int i = 0; i < count; ++i)
    if(problems[i].m > 0 && problems[i].n > 0)
      tiles[level[i]] += ceil_div(problems[i].m, cfg_bm(cfg[i])) * ceil_div(problems[i].n, cfg_bn(cfg[i]));
  for(int i = 0; i < count; ++i)
    if(level[i] > 0 && tiles[level[i]] < cus) cfg[i] = small;
  return cfg;
}

// Estimated duration (arbitrary units) of a launch under a policy: per level,
// the padded tile work spread over the CUs plus half a tile of tail.
double policy_estimate(const GroupedPolicy& pol, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const std::vector<int>& level, int cus) {
  const int levels = count ? *std::max_element(level.begin(), level.end()) + 1 : 0;
  std::vector<double> work(levels, 0.0), tmax(levels, 0.0);
  std::vector<int64_t> ntiles(levels, 0);
  const std::vector<int> cfg = assign_configs(pol, prec, problems, count, level, cus);
  for(int i = 0; i < count; ++i) {
    const auto& q = problems[i];
    if(q.m <= 0 || q.n <= 0) continue;
    const int c = cfg[i];
    const int64_t tiles = ceil_div(q.m, cfg_bm(c)) * ceil_div(q.n, cfg_bn(c));
    const double t = (double)cfg_bm(c) * cfg_bn(c) * (double)round_up(std::max<int64_t>(q.k, 1), 32) / kTileRate[c];
    work[level[i]] += t * tiles;
    tmax[level[i]] = std::max(tmax[level[i]], t);
    ntiles[level[i]] += tiles;
  }
  double est = 0.0;
  for(int l = 0; l < levels; ++l) {
    if(ntiles[l] == 0) continue;
    est += std::max(work[l] / cus + 0.5 * work[l] / ntiles[l], tmax[l]);
  }
  return est;
}

}  // namespace

int grouped_policy_count() { return kPolicyCount; }

bool grouped_policy_direct(int id) { return id >= 0 && id < kPolicyCount && kPolicies[id].direct; }

bool grouped_policy_info(int id, int* big, int* small, double* bound, int* per_level) {
  if(id < 0 || id >= kPolicyCount) return false;
  *big = kPolicies[id].big;
  *small = kPolicies[id].small;
  *bound = kPolicies[id].bound;
  *per_level = kPolicies[id].per_level;
  return true;
}

// Tuning-table key of a grouped launch: precision, problem count, total
// contraction volume and dependency depth, each bucketed by powers of two.
std::string grouped_tune_key(mfx_precision p, const mfx_gemm_problem* problems, int count, const mfx_dep* deps,
    int ndeps) {
  double vol = 0.0;
  for(int i = 0; i < count; ++i) vol += (double)problems[i].m * problems[i].n * problems[i].k;
  const auto level = problem_levels(count, deps, ndeps);
  const int levels = count ? *std::max_element(level.begin(), level.end()) + 1 : 0;
  return "g" + tune_key(p, count, (int64_t)std::min(std::sqrt(vol), 1e18), levels);
}

// Policy of a grouped launch: the tuner's override, a forced configuration,
// a tuned entry, or the estimate.
int choose_grouped_policy(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps) {
  if(h->grouped_policy >= 0) return h->grouped_policy;
  int forced = h->forced_config[prec];
  if(forced >= 0) {
    for(int i = 0; i < kPolicyCount; ++i)
      if(kPolicies[i].big == forced && kPolicies[i].small == forced && !kPolicies[i].per_level) return i;
  }
  const int tuned = tuned_lookup(grouped_tune_key(prec, problems, count, deps, ndeps));
  if(tuned >= 0 && tuned < kPolicyCount) return tuned;
  const auto level = problem_levels(count, deps, ndeps);
  const int cus = h->cu_count > 0 ? h->cu_count : 256;
  const bool levels = count > 0 && *std::max_element(level.begin(), level.end()) > 0;
  // With dependency levels: one launch per level, unless (BF16) at least
  // half of the work is in problems with neither producers nor consumers --
  // then one kernel with in-kernel waits, whose list-schedule order fills
  // the later levels' idle CUs with that work.  Measured on MI355X: a lone
  // 8-long 2048^2 chain 0.51 ms per level against 0.66 ms in one kernel;
  // the same chain beside 512 uneven problems 1.61 against 1.55 ms.
  bool one_kernel = !levels;
  if(levels && prec != MFX_PREC_FP32) {
    std::vector<char> linked(count, 0);
    for(int d = 0; d < ndeps; ++d) linked[deps[d].consumer] = linked[deps[d].producer] = 1;
    double free_work = 0.0, all_work = 0.0;
    for(int i = 0; i < count; ++i) {
      const double w = (double)problems[i].m * problems[i].n * problems[i].k;
      all_work += w;
      if(!linked[i]) free_work += w;
    }
    one_kernel = free_work >= 0.5 * all_work;
  }
  int best = 0;
  double best_est = -1.0;
  for(int i = 0; i < kPolicyCount; ++i) {
    // without levels the two modes are the same launch
    if(kPolicies[i].per_level == 2 || one_kernel != (kPolicies[i].per_level == 0)) continue;
    // BF16: the direct tiles (no packing pass; measured 1.3x faster on the
    // batched workload); FP32 has nothing to pack
    if(prec != MFX_PREC_FP32 && !kPolicies[i].direct) continue;
    const double e = policy_estimate(kPolicies[i], prec, problems, count, level, cus);
    if(best_est < 0 || e < best_est) {
      best_est = e;
      best = i;
    }
  }
  return best;
}

namespace {

// Clears the queue counters and the error flag before each launch (a kernel,
// so the reset can be part of a captured graph).
__global__ void zero_ints_kernel(int* p, int n, int* err) {
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
  if(blockIdx.x == 0 && threadIdx.x == 0) *err = 0;
}

// Bytes of the packed A and B copies of one problem (each rounded to 256).
std::pair<size_t, size_t> packed_sizes(const GProblem& P, int nhl) {
  const size_t ba = (size_t)round_up(P.M, 32) * P.kbs * 16 * nhl * 2;
  const size_t bb = (size_t)round_up(P.N, 32) * P.kbs * 16 * nhl * 2;
  return {(size_t)round_up(ba, 256), (size_t)round_up(bb, 256)};
}

}  // namespace

static mfx_status validate(mfx_precision prec, const mfx_gemm_problem* problems, int count, const mfx_dep* deps,
    int ndeps) {
  if(count < 0 || ndeps < 0) return invalid("negative problem or dependency count");
  if(count > 0 && !problems) return invalid("null problem list");
  if(ndeps > 0 && !deps) return invalid("null dependency list");
  for(int i = 0; i < count; ++i) {
    const auto& q = problems[i];
    if(q.m < 0 || q.n < 0 || q.k < 0 || q.m > (1 << 30) || q.n > (1 << 30) || q.k > (1 << 30))
      return invalid("problem " + std::to_string(i) + ": bad dimensions");
    if((q.transA != 0 && q.transA != 1) || (q.transB != 0 && q.transB != 1))
      return invalid("problem " + std::to_string(i) + ": transA / transB must be 0 or 1");
    const int64_t a_cols = q.transA ? q.m : q.k;
    const int64_t b_cols = q.transB ? q.k : q.n;
    if(q.m > 0 && q.n > 0) {
      if(prec != MFX_PREC_FP32 && !bf16_k_supported(q.k, prec == MFX_PREC_SPLIT_BF16 ? 2 : 1))
        return unsupported("problem " + std::to_string(i) + ": K is beyond the packed BF16 path's limit");
      if(!q.C || q.ldc < std::max<int64_t>(1, q.n)) return invalid("problem " + std::to_string(i) + ": bad C");
      if(q.k > 0 && (!q.A || !q.B)) return invalid("problem " + std::to_string(i) + ": null operand");
      if(q.lda < std::max<int64_t>(1, a_cols) || q.ldb < std::max<int64_t>(1, b_cols))
        return invalid("problem " + std::to_string(i) + ": leading dimension smaller than row length");
    }
  }
  for(int d = 0; d < ndeps; ++d) {
    if(deps[d].consumer < 0 || deps[d].consumer >= count || deps[d].producer < 0 || deps[d].producer >= count)
      return invalid("dependency " + std::to_string(d) + " names a problem out of range");
    if(deps[d].producer >= deps[d].consumer)
      return invalid("dependency " + std::to_string(d) + ": producers must come before their consumers");
  }
  return MFX_SUCCESS;
}

namespace {

// Cost model of the pipelined planner, in milliseconds, fitted to MI355X
// measurements of split-BF16 grouped launches: a tile of configuration c
// takes about kTileFixedMs + kTileMsPerUnit * bm * bn * K / kTileRate[c]
// (the 256x256 eight-wave tile at K = 2048: 0.2 ms under full load, with
// its panels streamed from HBM; the 512 uneven problems: 1.0 ms); the
// pre-pass packs kPrepassMsPerElem per operand element over the whole chip
// (0.53 ms for 361 M elements), and a pre-pack item costs one CU about
// kPrepackCuMsPerElem per element (the CU reads 4 and writes 4 bytes); and
// a launch moves at most kHbmBytesPerMs.
constexpr double kTileFixedMs = 0.005, kTileMsPerUnit = 1.67e-9;
constexpr double kPrepassMsPerElem = 1.47e-9, kPrepackCuMsPerElem = 1.0e-7, kHbmBytesPerMs = 5.0e9;
constexpr int kMinPipelineSegments = 3;
constexpr int PREPACK_BLOCKS = 32;  // 256-chunk blocks per pre-pack item (256 KiB of FP32 source)

double tile_ms(int c, int64_t k) {
  return kTileFixedMs + kTileMsPerUnit * cfg_bm(c) * cfg_bn(c) * (double)round_up(std::max<int64_t>(k, 1), 32) /
                            kTileRate[c];
}

// Pipelined segments.  Problems with producers or consumers keep their
// dependency level; every other problem goes, largest first, to the segment
// where it adds least to the estimated total time.  A segment's launch takes
// the longest of: its tile work plus the pre-pack items it runs for the next
// segment, spread over the CUs; its longest tile; its HBM traffic (each
// problem's packed operands read once -- its tiles share one XCD's L2 -- and
// C written, plus 4 + 2 * planes bytes per pre-packed element).  Its
// pre-pass comes on top: segment 0's packs all of its problems, a later
// one only those with producers; the others were packed by the previous
// segment's launch.
std::vector<int> balance_segments(const std::vector<GProblem>& P, const std::vector<char>& has_prod,
    const std::vector<char>& linked, const std::vector<int>& level, int cus, int nhl, int* nseg_out) {
  const bool packed = nhl > 0;
  const int count = (int)P.size();
  int nseg = 1;
  for(int i = 0; i < count; ++i) nseg = std::max(nseg, level[i] + 1);
  bool any_free = false;
  for(int i = 0; i < count; ++i) any_free = any_free || (!linked[i] && P[i].tiles_m * P[i].tiles_n > 0);
  if(any_free) nseg = std::max(nseg, kMinPipelineSegments);
  std::vector<double> work(nseg, 0.0), tmax(nseg, 0.0), prepass(nseg, 0.0), prepack(nseg, 0.0), bytes(nseg, 0.0);
  std::vector<int> seg(count, 0);
  auto elems = [&](int i) { return packed ? (double)(round_up(P[i].M, 32) + round_up(P[i].N, 32)) * P[i].kbs * 16 : 0.0; };
  auto add = [&](int i, int s, double sign) {
    const int tiles = P[i].tiles_m * P[i].tiles_n;
    if(tiles == 0) return;
    const double t = tile_ms(P[i].config, P[i].K);
    work[s] += sign * t * tiles;
    if(sign > 0) tmax[s] = std::max(tmax[s], t);
    const double e = elems(i);
    bytes[s] += sign * (e * 2 * std::max(nhl, 2) + (P[i].beta != 0.0f ? 8.0 : 4.0) * P[i].M * P[i].N);
    if(s == 0 || has_prod[i]) {
      prepass[s] += sign * e * kPrepassMsPerElem;
    } else {
      prepack[s - 1] += sign * e * kPrepackCuMsPerElem;
      bytes[s - 1] += sign * e * (4 + 2 * nhl);
    }
  };
  auto total = [&]() {
    double t = 0.0;
    for(int s = 0; s < nseg; ++s)
      t += std::max(std::max((work[s] + prepack[s]) / cus, tmax[s]), bytes[s] / kHbmBytesPerMs) + prepass[s];
    return t;
  };
  std::vector<int> free_list;
  for(int i = 0; i < count; ++i) {
    if(linked[i]) {
      seg[i] = level[i];
      add(i, seg[i], 1.0);
    } else if(P[i].tiles_m * P[i].tiles_n > 0) {
      free_list.push_back(i);
    }
  }
  std::stable_sort(free_list.begin(), free_list.end(), [&](int a, int b) {
    return tile_ms(P[a].config, P[a].K) * P[a].tiles_m * P[a].tiles_n >
           tile_ms(P[b].config, P[b].K) * P[b].tiles_m * P[b].tiles_n;
  });
  for(int i : free_list) {
    int best = 0;
    double best_t = -1.0;
    for(int s = 0; s < nseg; ++s) {
      const std::vector<double> keep = tmax;
      add(i, s, 1.0);
      const double t = total();
      add(i, s, -1.0);
      tmax = keep;
      if(best_t < 0 || t < best_t - 1e-12) {
        best_t = t;
        best = s;
      }
    }
    seg[i] = best;
    add(i, best, 1.0);
  }
  *nseg_out = nseg;
  return seg;
}

// Queue order of the one-kernel (in-kernel waits) mode when no problem has
// pack items: a list schedule of every tile on `workers` workgroups.  The
// next free workgroup takes a tile of the ready problem with the longest
// remaining path to the end of the launch (its tile time plus its longest
// consumer's, tile_ms estimates), ties to the larger K; a problem is ready
// once its producers' tiles are scheduled to have ended.  So a chain's
// later levels are queued as soon as their inputs are due, and the
// independent problems fill the time around them, instead of every later
// level waiting behind all first-level tiles.  A consumer's tiles follow all
// of its producers' tiles in the list, so the order stays a topological
// one: a workgroup that waits in-kernel only waits for tiles that other
// workgroups have already taken.  Producers have lower indices than their
// consumers (checked by validate).
std::vector<GItem> critical_path_items(const std::vector<int>& at, const std::vector<GProblem>& P, int workers) {
  const int count = (int)P.size();
  std::vector<char> in(count, 0);
  for(int i : at) in[i] = 1;
  std::vector<std::vector<int>> cons(count);
  std::vector<int> wait_for(count, 0);
  for(int i : at)
    for(int e = 0; e < P[i].ndeps; ++e)
      if(in[P[i].deps[e]]) {
        cons[P[i].deps[e]].push_back(i);
        ++wait_for[i];
      }
  std::vector<double> dur(count, 0.0), path(count, 0.0);
  for(int i : at) dur[i] = tile_ms(P[i].config, P[i].K);
  for(int i = count - 1; i >= 0; --i) {
    if(!in[i]) continue;
    double longest = 0.0;
    for(int c : cons[i]) longest = std::max(longest, path[c]);
    path[i] = dur[i] + longest;
  }
  auto lower = [&](int a, int b) {  // a goes after b
    if(path[a] != path[b]) return path[a] < path[b];
    if(P[a].K != P[b].K) return P[a].K < P[b].K;
    return a > b;
  };
  std::priority_queue<int, std::vector<int>, decltype(lower)> ready(lower);
  using Due = std::pair<double, int>;  // (time its inputs are due, problem)
  std::priority_queue<Due, std::vector<Due>, std::greater<Due>> pending;
  std::vector<double> due(count, 0.0), finish(count, 0.0);
  std::vector<int> next(count, 0);
  for(int i : at)
    if(wait_for[i] == 0) pending.push({0.0, i});
  std::priority_queue<double, std::vector<double>, std::greater<double>> free_at;
  for(int w = 0; w < std::max(1, workers); ++w) free_at.push(0.0);
  std::vector<GItem> out;
  while(!ready.empty() || !pending.empty()) {
    double t = free_at.top();
    free_at.pop();
    while(!pending.empty() && pending.top().first <= t) {
      ready.push(pending.top().second);
      pending.pop();
    }
    if(ready.empty()) {  // nothing due yet: this workgroup waits for the earliest
      t = pending.top().first;
      ready.push(pending.top().second);
      pending.pop();
    }
    const int p = ready.top();
    const int tile = next[p]++;
    out.push_back({ITEM_TILE, p, tile / P[p].tiles_n, tile % P[p].tiles_n});
    finish[p] = std::max(finish[p], t + dur[p]);
    free_at.push(t + dur[p]);
    if(next[p] == P[p].tiles_m * P[p].tiles_n) {
      ready.pop();
      for(int c : cons[p]) {
        due[c] = std::max(due[c], finish[p]);
        if(--wait_for[c] == 0) pending.push({due[c], c});
      }
    }
  }
  return out;
}

// Tiles of one problem that go to one queue: at most one round of an XCD's
// CUs (kQueueGroupTiles), whole tile rows where possible, so the group's
// tiles share their A panels (and B panels) in that XCD's L2.
constexpr int kQueueGroupTiles = 32;
struct TileGroup {
  int problem, first, count;  // tiles [first, first + count) in row-major tile order
};

// Splits a problem list over the GROUPED_QUEUES queues: each problem's tiles
// in groups of at most kQueueGroupTiles, groups by decreasing work to the
// least-loaded queue; a queue lists its groups in the order of `order`
// (largest K first), so long tiles start first everywhere.
std::vector<std::vector<TileGroup>> queue_tiles(const std::vector<int>& order, const std::vector<GProblem>& P) {
  std::vector<TileGroup> groups;
  std::vector<int> rank(P.size(), 0);
  for(size_t r = 0; r < order.size(); ++r) rank[order[r]] = (int)r;
  for(int i : order) {
    const int tiles = P[i].tiles_m * P[i].tiles_n;
    const int per = std::min(tiles, kQueueGroupTiles);
    const int ng = (tiles + per - 1) / per;
    for(int g = 0; g < ng; ++g) {
      const int first = (int)((int64_t)tiles * g / ng), last = (int)((int64_t)tiles * (g + 1) / ng);
      groups.push_back({i, first, last - first});
    }
  }
  auto work = [&](const TileGroup& g) { return tile_ms(P[g.problem].config, P[g.problem].K) * g.count; };
  std::vector<int> by_work(groups.size());
  std::iota(by_work.begin(), by_work.end(), 0);
  std::stable_sort(by_work.begin(), by_work.end(), [&](int a, int b) { return work(groups[a]) > work(groups[b]); });
  std::vector<double> load(GROUPED_QUEUES, 0.0);
  std::vector<int> q_of(groups.size(), 0);
  for(int gi : by_work) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    q_of[gi] = q;
    load[q] += work(groups[gi]);
  }
  std::vector<std::vector<TileGroup>> qs(GROUPED_QUEUES);
  for(size_t gi = 0; gi < groups.size(); ++gi) qs[q_of[gi]].push_back(groups[gi]);  // groups are in `order`
  return qs;
}

PackJob make_job(const GProblem& P, bool is_a, int first_block) {
  PackJob j;
  j.src = make_pack_src(is_a ? P.A : P.B, is_a ? P.lda : P.ldb, is_a ? P.M : P.N, P.K, is_a ? P.a_kmajor : P.b_kmajor);
  j.dst = nullptr;  // bound to the workspace by bind_grouped
  j.kbs = P.kbs;
  j.chunks = ceil_div(is_a ? P.M : P.N, 32) * P.kbs * 64;
  j.blocks = (int)ceil_div(j.chunks, 256);
  j.first_block = first_block;
  return j;
}

}  // namespace

mfx_status plan_grouped(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps, GroupedPlan& pl) {
  if(prec != MFX_PREC_FP32 && prec != MFX_PREC_SPLIT_BF16 && prec != MFX_PREC_BF16)
    return unsupported("grouped launch supports FP32, SPLIT_BF16 and BF16");
  MFX_CHECK(validate(prec, problems, count, deps, ndeps));
  pl.prec = prec;
  pl.problems.resize(count);
  std::vector<int> level(count, 0);
  for(int i = 0; i < count; ++i) pl.problems[i].ndeps = pl.problems[i].consumed = 0;
  for(int d = 0; d < ndeps; ++d) {
    GProblem& c = pl.problems[deps[d].consumer];
    bool dup = false;
    for(int e = 0; e < c.ndeps; ++e) dup |= c.deps[e] == deps[d].producer;
    if(dup) continue;
    if(c.ndeps == GROUPED_MAX_DEPS) return unsupported("a problem may depend on at most 4 producers");
    c.deps[c.ndeps++] = deps[d].producer;
    pl.problems[deps[d].producer].consumed = 1;
  }
  for(int i = 0; i < count; ++i)
    for(int e = 0; e < pl.problems[i].ndeps; ++e) level[i] = std::max(level[i], level[pl.problems[i].deps[e]] + 1);
  std::vector<char> has_prod(count), linked(count);
  for(int i = 0; i < count; ++i) {
    has_prod[i] = pl.problems[i].ndeps > 0;
    linked[i] = has_prod[i] || pl.problems[i].consumed;
  }
  pl.policy = choose_grouped_policy(h, prec, problems, count, deps, ndeps);
  const GroupedPolicy& pol = kPolicies[pl.policy];
  pl.big = grouped_config(prec, pol.big);
  pl.small = grouped_config(prec, pol.small);
  const int cus = h->cu_count > 0 ? h->cu_count : 256;
  const bool packed = prec != MFX_PREC_FP32;
  const int nhl = prec == MFX_PREC_SPLIT_BF16 ? 2 : 1;
  const int max_level = count ? *std::max_element(level.begin(), level.end()) : 0;
  pl.pipelined = pol.per_level == 2;
  // One launch per segment: kernel boundaries order the segments, so no
  // problem waits or publishes in-kernel and every operand is packed by a
  // pre-pass or by the previous segment's pre-pack items.
  pl.per_level = (pol.per_level == 1 && max_level > 0) || pl.pipelined;
  std::vector<int> cfg(count);
  if(pl.pipelined) {
    // Problems with producers sit on the critical path of the segments (a
    // chain's later levels): they take the small tile, so each of their
    // segments is short and the independent problems fill it.
    for(int i = 0; i < count; ++i)
      cfg[i] = has_prod[i] ? pl.small : problem_config(pol, prec, problems[i].m, problems[i].n);
  } else {
    cfg = assign_configs(pol, prec, problems, count, level, cus);
  }
  // With a launch per level, a later level whose tiles leave CUs idle (the
  // last products of a chain) may take a smaller tile of the same workgroup
  // size for all its problems, in a kernel of its own tile pair, when that
  // is the least estimated time in rounds over the CUs.
  std::vector<int> level_cfg(max_level + 1, -1);
  if(pl.per_level && !pl.pipelined) {
    const int threads = cfg_threads(pl.big);
    for(int lv = 1; lv <= max_level; ++lv) {
      auto level_tiles = [&](int c) {
        int64_t tiles = 0;
        for(int i = 0; i < count; ++i)
          if(level[i] == lv && problems[i].m > 0 && problems[i].n > 0)
            tiles += ceil_div(problems[i].m, cfg_bm(c >= 0 ? c : cfg[i])) * ceil_div(problems[i].n, cfg_bn(c >= 0 ? c : cfg[i]));
        return tiles;
      };
      if(level_tiles(-1) >= cus) continue;
      double best_t = -1.0;
      int best = -1;
      for(int c = 0; c < 7; ++c) {  // tile configurations 0..6
        if(grouped_config(prec, c) != c || cfg_threads(c) != threads) continue;
        const double t = (double)ceil_div(level_tiles(c), cus) * cfg_bm(c) * cfg_bn(c) / config_rate(prec, c);
        if(best < 0 || t < best_t) {
          best_t = t;
          best = c;
        }
      }
      bool uniform = true;
      for(int i = 0; i < count; ++i) uniform = uniform && (level[i] != lv || cfg[i] == best);
      if(best >= 0 && !uniform) {
        level_cfg[lv] = best;
        for(int i = 0; i < count; ++i)
          if(level[i] == lv) cfg[i] = best;
      }
    }
  }

  size_t packed_bytes = 0;
  int panels = 0;
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
    P.config = cfg[i];
    const int bm = cfg_bm(P.config), bn = cfg_bn(P.config);
    P.tiles_m = (int)ceil_div(q.m, bm);
    P.tiles_n = (int)ceil_div(q.n, bn);
    if(q.m == 0 || q.n == 0) P.tiles_m = P.tiles_n = 0;
    P.vec_a = f32_vec_ok(q.A, q.lda, P.a_kmajor, P.M, P.K);
    P.vec_b = f32_vec_ok(q.B, q.ldb, P.b_kmajor, P.N, P.K);
    if(P.tiles_m * P.tiles_n > 0) pl.all_vec = pl.all_vec && P.vec_a && P.vec_b;
    P.kbs = 0;
    P.pack_slices = 0;
    P.packed_a = P.packed_b = nullptr;
    P.pack_a_ctr = P.pack_b_ctr = 0;
    P.prepacked = 0;
    P.direct = packed && pol.direct && P.tiles_m * P.tiles_n > 0 && q.k % 16 == 0 &&
               direct_operand_ok(q.A, q.lda, P.M, P.a_kmajor) && direct_operand_ok(q.B, q.ldb, P.N, P.b_kmajor);
    if(P.direct) {
      P.prepacked = 1;  // nothing to pack: the tiles read A and B themselves
    } else if(packed && P.tiles_m * P.tiles_n > 0) {
      // load_chunk8 takes 128-bit loads along k only
      P.vec_a = P.a_kmajor && P.vec_a;
      P.vec_b = P.b_kmajor && P.vec_b;
      P.kbs = (int)(round_up(q.k, 16 * kb_per_stage_of(nhl)) / 16);
      P.pack_slices = (int)ceil_div(P.kbs, PACK_KBLOCKS);
      packed_bytes += packed_sizes(P, nhl).first + packed_sizes(P, nhl).second;
      if(P.ndeps == 0 || pl.per_level) {
        // packed by a pre-pass kernel or pre-pack items before the tiles start
        P.prepacked = 1;
      } else {
        P.pack_a_ctr = panels;
        panels += P.tiles_m;
        P.pack_b_ctr = panels;
        panels += P.tiles_n;
      }
    }
    pl.flops += 2.0 * q.m * q.n * q.k;
  }

  // Segment of every problem.
  int nseg = 1;
  std::vector<int> seg(count, 0);
  if(pl.pipelined) {
    seg = balance_segments(pl.problems, has_prod, linked, level, cus, packed ? nhl : 0, &nseg);
  } else if(pl.per_level) {
    seg = level;
    nseg = max_level + 1;
  }

  // Items segment by segment.  In-kernel mode (one segment): the level's
  // pack items (largest K first, so the first tiles' panels are ready first),
  // then its tiles, largest K first, a problem's tiles row by row, all in
  // queue 0.  Launch-per-segment modes: each problem's tiles in one XCD
  // queue, problems largest K first, with the segment's pre-pack items
  // spread evenly over the queues.
  std::vector<std::pair<int, int>> pre_ops, prepack_ops;  // (problem, is_a)
  std::vector<int> pre_seg_end, prepack_seg_end;
  pl.seg_items.assign(1, 0);
  for(int sg = 0; sg < nseg; ++sg) {
    std::vector<int> at;
    for(int i = 0; i < count; ++i)
      if(seg[i] == sg && pl.problems[i].tiles_m * pl.problems[i].tiles_n > 0) at.push_back(i);
    std::stable_sort(at.begin(), at.end(), [&](int a, int b) { return pl.problems[a].K > pl.problems[b].K; });
    // operands packed by this segment's pre-pass
    for(int i : at)
      if(pl.problems[i].prepacked && pl.problems[i].kbs > 0 && (!pl.pipelined || sg == 0 || has_prod[i])) {
        pre_ops.push_back({i, 1});
        pre_ops.push_back({i, 0});
      }
    pre_seg_end.push_back((int)pre_ops.size());
    // operands of the next segment packed by this segment's pre-pack items
    const int prepack_first = (int)prepack_ops.size();
    if(pl.pipelined && packed && sg + 1 < nseg)
      for(int i = 0; i < count; ++i)
        if(seg[i] == sg + 1 && !has_prod[i] && pl.problems[i].kbs > 0 && pl.problems[i].tiles_m * pl.problems[i].tiles_n > 0) {
          prepack_ops.push_back({i, 1});
          prepack_ops.push_back({i, 0});
        }
    prepack_seg_end.push_back((int)prepack_ops.size());
    const size_t first_item = pl.items.size();
    bool no_pack_items = true;
    for(int i : at) no_pack_items = no_pack_items && (!packed || pl.problems[i].prepacked);
    if(!pl.per_level && no_pack_items) {
      // one queue in list-schedule order
      const auto order = critical_path_items(at, pl.problems, cus);
      pl.items.insert(pl.items.end(), order.begin(), order.end());
      pl.qbeg.push_back(0);
      for(int q = 0; q < GROUPED_QUEUES; ++q) pl.qbeg.push_back((int)(pl.items.size() - first_item));
    } else if(!pl.per_level) {
      // one queue, dependency level by level (pack items before their level's tiles)
      for(int lv = 0; lv <= max_level; ++lv) {
        if(packed)
          for(int i : at) {
            const GProblem& P = pl.problems[i];
            if(level[i] != lv || P.prepacked) continue;
            for(int b = 0; b < (int)ceil_div(P.M, 32); ++b)
              for(int sl = 0; sl < P.pack_slices; ++sl) pl.items.push_back({ITEM_PACK_A, i, b, sl});
            for(int b = 0; b < (int)ceil_div(P.N, 32); ++b)
              for(int sl = 0; sl < P.pack_slices; ++sl) pl.items.push_back({ITEM_PACK_B, i, b, sl});
          }
        for(int i : at)
          if(level[i] == lv)
            for(int tm = 0; tm < pl.problems[i].tiles_m; ++tm)
              for(int tn = 0; tn < pl.problems[i].tiles_n; ++tn) pl.items.push_back({ITEM_TILE, i, tm, tn});
      }
      pl.qbeg.push_back(0);
      for(int q = 0; q < GROUPED_QUEUES; ++q) pl.qbeg.push_back((int)(pl.items.size() - first_item));
    } else {
      // pre-pack items: PREPACK_BLOCKS blocks of this segment's pre-pack job list each
      int blocks = 0;
      for(int j = prepack_first; j < (int)prepack_ops.size(); ++j) {
        const GProblem& P = pl.problems[prepack_ops[j].first];
        blocks += (int)ceil_div(ceil_div(prepack_ops[j].second ? P.M : P.N, 32) * P.kbs * 64, 256);
      }
      std::vector<GItem> pre_items;
      for(int b = 0; b < blocks; b += PREPACK_BLOCKS)
        pre_items.push_back({ITEM_PREPACK, 0, b, std::min(PREPACK_BLOCKS, blocks - b)});
      const auto qs = queue_tiles(at, pl.problems);
      pl.qbeg.push_back(0);
      size_t next_pre = 0;
      for(int q = 0; q < GROUPED_QUEUES; ++q) {
        std::vector<GItem> tiles;
        for(const TileGroup& tg : qs[q]) {
          const int tn_count = pl.problems[tg.problem].tiles_n;
          for(int t = tg.first; t < tg.first + tg.count; ++t)
            tiles.push_back({ITEM_TILE, tg.problem, t / tn_count, t % tn_count});
        }
        // this queue's share of the pre-pack items, spread evenly over its tiles
        const size_t share = (pre_items.size() * (q + 1)) / GROUPED_QUEUES - (pre_items.size() * q) / GROUPED_QUEUES;
        const size_t nt = tiles.size();
        size_t ti = 0;
        for(size_t k = 0; k < share; ++k) {
          const size_t upto = nt * (k + 1) / (share + 1);
          while(ti < upto) pl.items.push_back(tiles[ti++]);
          pl.items.push_back(pre_items[next_pre++]);
        }
        while(ti < nt) pl.items.push_back(tiles[ti++]);
        pl.qbeg.push_back((int)(pl.items.size() - first_item));
      }
    }
    pl.seg_items.push_back((int)pl.items.size());
    pl.seg_big.push_back(pl.per_level && !pl.pipelined && level_cfg[sg] >= 0 ? level_cfg[sg] : pl.big);
    pl.seg_small.push_back(pl.per_level && !pl.pipelined && level_cfg[sg] >= 0 ? level_cfg[sg] : pl.small);
  }
  if(pl.per_level)
    for(auto& P : pl.problems) P.ndeps = P.consumed = 0;  // ordered by the launches
  const int segments = nseg;

  // Job tables: pre-pass jobs (block numbers from 0 within each segment),
  // then pre-pack jobs (likewise).
  pl.jobs.clear();
  pl.job_op.clear();
  pl.seg_jobs.assign(1, 0);
  pl.seg_blocks.assign(segments, 0);
  for(int sg = 0, j = 0; sg < segments; ++sg) {
    int blocks = 0;
    for(; j < pre_seg_end[sg]; ++j) {
      const PackJob job = make_job(pl.problems[pre_ops[j].first], pre_ops[j].second == 1, blocks);
      blocks += job.blocks;
      pl.jobs.push_back(job);
      pl.job_op.push_back(pre_ops[j]);
    }
    pl.seg_blocks[sg] = blocks;
    pl.seg_jobs.push_back((int)pl.jobs.size());
  }
  pl.npre = (int)pl.jobs.size();
  pl.seg_prejobs.assign(1, pl.npre);
  for(int sg = 0, j = 0; sg < segments; ++sg) {
    int blocks = 0;
    for(; j < prepack_seg_end[sg]; ++j) {
      const PackJob job = make_job(pl.problems[prepack_ops[j].first], prepack_ops[j].second == 1, blocks);
      blocks += job.blocks;
      pl.jobs.push_back(job);
      pl.job_op.push_back(prepack_ops[j]);
    }
    pl.seg_prejobs.push_back((int)pl.jobs.size());
  }

  pl.off_problems = 0;
  pl.off_items = round_up(sizeof(GProblem) * std::max(1, count), 256);
  pl.off_qbeg = pl.off_items + round_up(sizeof(GItem) * std::max<size_t>(1, pl.items.size()), 256);
  pl.off_jobs = pl.off_qbeg + round_up(sizeof(int) * pl.qbeg.size(), 256);
  pl.off_counters = pl.off_jobs + round_up(sizeof(PackJob) * std::max<size_t>(1, pl.jobs.size()), 256);
  // counters: error flag, GROUPED_QUEUES queue heads per segment, per-problem done, per-panel pack counts
  pl.counter_ints = 1 + (size_t)segments * GROUPED_QUEUES + (size_t)count + (size_t)panels;
  pl.off_packed = pl.off_counters + round_up(pl.counter_ints * sizeof(int), 256);
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
      if(P.kbs == 0) continue;
      const auto sz = packed_sizes(P, nhl);
      P.packed_a = reinterpret_cast<u32x4*>(base + off);
      off += sz.first;
      P.packed_b = reinterpret_cast<u32x4*>(base + off);
      off += sz.second;
    }
    for(size_t j = 0; j < pl.jobs.size(); ++j) {
      const GProblem& P = pl.problems[pl.job_op[j].first];
      pl.jobs[j].dst = pl.job_op[j].second == 1 ? P.packed_a : P.packed_b;
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
  memcpy(st + pl.off_qbeg, pl.qbeg.data(), sizeof(int) * pl.qbeg.size());
  if(!pl.jobs.empty()) memcpy(st + pl.off_jobs, pl.jobs.data(), sizeof(PackJob) * pl.jobs.size());
  MFX_HIP(hipMemcpyAsync(base, st, pl.off_counters, hipMemcpyHostToDevice, s));
  MFX_HIP(hipEventRecord(h->staging_ready, s));
  int* counters = reinterpret_cast<int*>(base + pl.off_counters);
  g.problems = reinterpret_cast<const GProblem*>(base + pl.off_problems);
  g.items = reinterpret_cast<const GItem*>(base + pl.off_items);
  g.qbeg = reinterpret_cast<const int*>(base + pl.off_qbeg);
  pl.dev_jobs = reinterpret_cast<const PackJob*>(base + pl.off_jobs);
  g.nitems = (int)pl.items.size();
  const int segments = (int)pl.seg_items.size() - 1;
  g.error = counters;
  g.head = counters + 1;  // segment 0; segment l uses head + l * GROUPED_QUEUES
  g.done = counters + 1 + (size_t)segments * GROUPED_QUEUES;
  g.pack_ctr = g.done + pl.problems.size();
  g.spin_limit = h->spin_limit;
  g.prejobs = nullptr;
  g.nprejobs = 0;
  return MFX_SUCCESS;
}

// Enqueues the counter reset, then per segment the pre-pass and the
// persistent kernel.
static mfx_status launch_grouped_kernels(Handle* h, const GroupedPlan& pl, const GroupedArgs& g, hipStream_t s) {
  const int n = (int)pl.counter_ints;
  MFX_CHECK(timed_launch(h, "grouped_reset", s, [&] {
    zero_ints_kernel<<<(int)std::min<int64_t>(ceil_div(n, 256), 1024), 256, 0, s>>>(g.head - 1, n, g.error);
  }));
  if(pl.items.empty()) return MFX_SUCCESS;
  const int grid = h->cu_count > 0 ? h->cu_count : 256;
  const int nhl = pl.prec == MFX_PREC_SPLIT_BF16 ? 2 : 1;
  const int segments = (int)pl.seg_items.size() - 1;
  for(int l = 0; l < segments; ++l) {
    if(pl.prec != MFX_PREC_FP32)
      MFX_CHECK(launch_pack_jobs(h, nhl, pl.dev_jobs + pl.seg_jobs[l], pl.seg_jobs[l + 1] - pl.seg_jobs[l],
          pl.seg_blocks[l], s));
    GroupedArgs gl = g;
    gl.items = g.items + pl.seg_items[l];
    gl.nitems = pl.seg_items[l + 1] - pl.seg_items[l];
    gl.head = g.head + (size_t)l * GROUPED_QUEUES;
    gl.qbeg = g.qbeg + (size_t)l * (GROUPED_QUEUES + 1);
    gl.prejobs = pl.dev_jobs + pl.seg_prejobs[l];
    gl.nprejobs = pl.seg_prejobs[l + 1] - pl.seg_prejobs[l];
    if(gl.nitems == 0) continue;
    const int big = pl.seg_big[l], small = pl.seg_small[l];
    if(pl.prec == MFX_PREC_FP32) MFX_CHECK(launch_grouped_f32(h, big, small, pl.all_vec, gl, grid, s));
    else MFX_CHECK(launch_grouped_bf16(h, nhl, big, small, pl.all_vec, gl, grid, s));
  }
  return MFX_SUCCESS;
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

mfx_status time_grouped(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps, int reps, hipStream_t s, double* ms) {
  GroupedPlan pl;
  MFX_CHECK(plan_grouped(h, prec, problems, count, deps, ndeps, pl));
  void* ws = nullptr;
  MFX_CHECK(workspace(h, pl.total_bytes, &ws));
  GroupedArgs g;
  MFX_CHECK(bind_grouped(h, pl, ws, g, s));
  if(!h->error_flag) MFX_HIP(hipMalloc(&h->error_flag, 256));
  g.error = h->error_flag;
  MFX_CHECK(launch_grouped_kernels(h, pl, g, s));  // untimed
  hipEvent_t a = nullptr, b = nullptr;
  MFX_HIP(hipEventCreate(&a));
  hipError_t e = hipEventCreate(&b);
  mfx_status st = e == hipSuccess ? MFX_SUCCESS : hip_fail(e, "hipEventCreate");
  if(st == MFX_SUCCESS && (e = hipEventRecord(a, s)) != hipSuccess) st = hip_fail(e, "hipEventRecord");
  for(int r = 0; r < reps && st == MFX_SUCCESS; ++r) st = launch_grouped_kernels(h, pl, g, s);
  float t = 0.0f;
  if(st == MFX_SUCCESS) {
    e = hipEventRecord(b, s);
    if(e == hipSuccess) e = hipEventSynchronize(b);
    if(e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if(e != hipSuccess) st = hip_fail(e, "grouped timing");
  }
  (void)hipEventDestroy(a);
  if(b) (void)hipEventDestroy(b);
  *ms = t / std::max(reps, 1);
  return st;
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
  void* mem = nullptr;  // tables, counters and packed operands of this graph
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
  // The graph owns its tables, counters and packed operands, so later calls on
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
