// Host-only check of the grouped-launch planner (mfx::plan_grouped): for
// many problem lists, every precision and every tile policy, the plan must
// cover every output tile exactly once, pack every operand exactly once and
// before its tiles run, respect every dependency, and keep its queue, job and
// block tables consistent.  No device is touched: planning only reads the
// problem descriptions (the operand pointers are stand-ins).
//
// Built and run by tests/test_grouped_plan.py; prints "OK <plans>" or the
// first violated invariant and exits non-zero.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "grouped.h"

using namespace mfx;

namespace mfx {
int grouped_policy_count();
bool grouped_policy_info(int id, int* big, int* small, double* bound, int* per_level);
}  // namespace mfx

static int g_failures = 0;
#define CHECK(cond, ...)                                   \
  do {                                                     \
    if(!(cond)) {                                          \
      std::fprintf(stderr, "FAIL %s: ", ctx.c_str());      \
      std::fprintf(stderr, __VA_ARGS__);                   \
      std::fprintf(stderr, "  [%s:%d]\n", __FILE__, __LINE__); \
      ++g_failures;                                        \
      return;                                              \
    }                                                      \
  } while(0)

struct Case {
  std::string name;
  std::vector<mfx_gemm_problem> probs;
  std::vector<mfx_dep> deps;
};

static float* stand_in(size_t i) { return reinterpret_cast<float*>((uintptr_t)(0x100000 + i * 0x1000)); }

static mfx_gemm_problem prob(int64_t m, int64_t n, int64_t k, size_t id, int ta = 0, int tb = 0) {
  mfx_gemm_problem p{};
  p.m = m;
  p.n = n;
  p.k = k;
  p.alpha = 1.0f;
  p.beta = 0.0f;
  p.A = stand_in(3 * id);
  p.lda = ta ? std::max<int64_t>(m, 1) : std::max<int64_t>(k, 1);
  p.transA = ta;
  p.B = stand_in(3 * id + 1);
  p.ldb = tb ? std::max<int64_t>(k, 1) : std::max<int64_t>(n, 1);
  p.transB = tb;
  p.C = stand_in(3 * id + 2);
  p.ldc = std::max<int64_t>(n, 1);
  return p;
}

static void add_chain(Case& c, int len, int64_t n) {
  std::vector<int> level;
  for(int i = 0; i < len; ++i) level.push_back(-1);
  std::vector<int> cur = level;
  while(cur.size() > 1) {
    std::vector<int> up;
    for(size_t i = 0; i + 1 < cur.size(); i += 2) {
      const int idx = (int)c.probs.size();
      c.probs.push_back(prob(n, n, n, idx));
      if(cur[i] >= 0) c.deps.push_back({idx, cur[i]});
      if(cur[i + 1] >= 0) c.deps.push_back({idx, cur[i + 1]});
      up.push_back(idx);
    }
    if(cur.size() % 2) up.push_back(cur.back());
    cur.swap(up);
  }
}

static void check_plan(const Case& cs, mfx_precision prec, int policy) {
  const std::string ctx = cs.name + " prec " + std::to_string((int)prec) + " policy " + std::to_string(policy);
  Handle h;
  h.cu_count = 256;
  h.grouped_policy = policy;
  GroupedPlan pl;
  const int count = (int)cs.probs.size();
  mfx_status st = plan_grouped(&h, prec, cs.probs.data(), count, cs.deps.data(), (int)cs.deps.size(), pl);
  CHECK(st == MFX_SUCCESS, "plan_grouped returned %d", (int)st);
  const bool packed = prec != MFX_PREC_FP32;
  const int segments = (int)pl.seg_items.size() - 1;
  CHECK(segments >= 1, "no segment");
  CHECK((int)pl.qbeg.size() == segments * (GROUPED_QUEUES + 1), "queue table size %zu", pl.qbeg.size());
  CHECK(pl.seg_items.front() == 0 && pl.seg_items.back() == (int)pl.items.size(), "segment item bounds");
  CHECK((int)pl.seg_jobs.size() == segments + 1 && (int)pl.seg_prejobs.size() == segments + 1, "segment job bounds");
  CHECK((int)pl.seg_blocks.size() == segments, "segment block counts");
  size_t panels = 0;
  for(const GProblem& P : pl.problems)
    if(packed && !P.prepacked && P.tiles_m * P.tiles_n > 0) panels += P.tiles_m + P.tiles_n;
  CHECK(pl.counter_ints == 1 + (size_t)segments * GROUPED_QUEUES + (size_t)count + panels, "counter layout");
  CHECK(pl.off_items >= sizeof(GProblem) * count && pl.off_qbeg > pl.off_items && pl.off_jobs > pl.off_qbeg &&
            pl.off_counters > pl.off_jobs && pl.off_packed > pl.off_counters && pl.total_bytes >= pl.off_packed,
      "table offsets");
  // queue bounds
  for(int l = 0; l < segments; ++l) {
    const int* q = &pl.qbeg[(size_t)l * (GROUPED_QUEUES + 1)];
    CHECK(q[0] == 0 && q[GROUPED_QUEUES] == pl.seg_items[l + 1] - pl.seg_items[l], "segment %d queue ends", l);
    for(int k = 0; k < GROUPED_QUEUES; ++k) CHECK(q[k] <= q[k + 1], "segment %d queue %d decreasing", l, k);
    if(!pl.per_level) CHECK(q[1] == q[GROUPED_QUEUES], "in-kernel waits need one queue");
  }
  // tiles: exactly once, one segment per problem
  std::vector<int> tseg(count, -1);
  std::map<std::tuple<int, int, int>, int> seen;
  std::vector<int> first_item(count, -1), last_item(count, -1);
  for(int l = 0; l < segments; ++l)
    for(int it = pl.seg_items[l]; it < pl.seg_items[l + 1]; ++it) {
      const GItem& g = pl.items[it];
      if(g.type == ITEM_TILE) {
        CHECK(g.problem >= 0 && g.problem < count, "tile of problem %d", g.problem);
        const GProblem& P = pl.problems[g.problem];
        CHECK(g.x >= 0 && g.x < P.tiles_m && g.y >= 0 && g.y < P.tiles_n, "tile (%d,%d) outside problem %d", g.x, g.y,
            g.problem);
        const int times = ++seen[std::make_tuple(g.problem, g.x, g.y)];
        CHECK(times == 1, "tile (%d,%d) of problem %d twice", g.x, g.y, g.problem);
        CHECK(tseg[g.problem] < 0 || tseg[g.problem] == l, "problem %d in two segments", g.problem);
        tseg[g.problem] = l;
        if(first_item[g.problem] < 0) first_item[g.problem] = it;
        last_item[g.problem] = it;
      } else if(g.type == ITEM_PREPACK) {
        CHECK(pl.per_level && packed, "pre-pack item outside a packed per-segment launch");
      } else {
        CHECK(!pl.per_level && packed, "in-kernel pack item in a per-segment launch");
      }
    }
  for(int i = 0; i < count; ++i) {
    const GProblem& P = pl.problems[i];
    const bool empty = cs.probs[i].m == 0 || cs.probs[i].n == 0;
    CHECK(P.tiles_m * P.tiles_n == (empty ? 0 : (int)(ceil_div(P.M, cfg_bm(P.config)) * ceil_div(P.N, cfg_bn(P.config)))),
        "problem %d tile count", i);
    int n = 0;
    for(auto& kv : seen) n += std::get<0>(kv.first) == i;
    CHECK(n == P.tiles_m * P.tiles_n, "problem %d: %d of %d tiles planned", i, n, P.tiles_m * P.tiles_n);
  }
  // dependencies
  for(const auto& d : cs.deps) {
    if(pl.problems[d.consumer].tiles_m * pl.problems[d.consumer].tiles_n == 0) continue;
    if(pl.problems[d.producer].tiles_m * pl.problems[d.producer].tiles_n == 0) continue;
    if(pl.per_level) {
      CHECK(tseg[d.producer] < tseg[d.consumer], "consumer %d in segment %d, producer %d in %d", d.consumer,
          tseg[d.consumer], d.producer, tseg[d.producer]);
    } else {
      CHECK(last_item[d.producer] < first_item[d.consumer], "consumer %d queued before producer %d", d.consumer,
          d.producer);
    }
  }
  if(!packed) {
    CHECK(pl.jobs.empty(), "FP32 plan packs");
    return;
  }
  // pack jobs: chunk and block counts, numbering within each segment's list
  CHECK(pl.jobs.size() == pl.job_op.size(), "job table sizes");
  for(size_t j = 0; j < pl.jobs.size(); ++j) {
    const PackJob& J = pl.jobs[j];
    const GProblem& P = pl.problems[pl.job_op[j].first];
    const int rows = pl.job_op[j].second ? P.M : P.N;
    CHECK(J.kbs == P.kbs && J.chunks == ceil_div(rows, 32) * P.kbs * 64 && J.blocks == (int)ceil_div(J.chunks, 256),
        "job %zu sizes", j);
  }
  auto check_blocks = [&](int j0, int j1, int* total) {
    int b = 0;
    for(int j = j0; j < j1; ++j) {
      if(pl.jobs[j].first_block != b) return false;
      b += pl.jobs[j].blocks;
    }
    *total = b;
    return true;
  };
  std::map<std::pair<int, int>, int> packs;  // (problem, is_a) -> count
  for(int l = 0; l < segments; ++l) {
    int blocks = 0;
    CHECK(check_blocks(pl.seg_jobs[l], pl.seg_jobs[l + 1], &blocks), "segment %d pre-pass block numbering", l);
    CHECK(blocks == pl.seg_blocks[l], "segment %d pre-pass block total", l);
    for(int j = pl.seg_jobs[l]; j < pl.seg_jobs[l + 1]; ++j) {
      ++packs[pl.job_op[j]];
      CHECK(tseg[pl.job_op[j].first] == l, "pre-pass of segment %d packs problem %d of segment %d", l,
          pl.job_op[j].first, tseg[pl.job_op[j].first]);
    }
    CHECK(check_blocks(pl.seg_prejobs[l], pl.seg_prejobs[l + 1], &blocks), "segment %d pre-pack block numbering", l);
    for(int j = pl.seg_prejobs[l]; j < pl.seg_prejobs[l + 1]; ++j) {
      const int i = pl.job_op[j].first;
      ++packs[pl.job_op[j]];
      CHECK(tseg[i] == l + 1, "segment %d pre-packs problem %d of segment %d", l, i, tseg[i]);
      for(const auto& d : cs.deps) CHECK(d.consumer != i, "segment %d pre-packs consumer %d", l, i);
    }
    // pre-pack items cover the segment's blocks exactly once
    std::vector<int> cover(blocks, 0);
    for(int it = pl.seg_items[l]; it < pl.seg_items[l + 1]; ++it) {
      const GItem& g = pl.items[it];
      if(g.type != ITEM_PREPACK) continue;
      CHECK(g.x >= 0 && g.y > 0 && g.x + g.y <= blocks, "segment %d pre-pack item [%d, %d) of %d", l, g.x, g.x + g.y,
          blocks);
      for(int b = g.x; b < g.x + g.y; ++b) ++cover[b];
    }
    for(int b = 0; b < blocks; ++b) CHECK(cover[b] == 1, "segment %d pre-pack block %d covered %d times", l, b, cover[b]);
  }
  for(int i = 0; i < count; ++i) {
    const GProblem& P = pl.problems[i];
    if(P.direct) {
      const auto& q = cs.probs[i];
      const bool ok = q.k % 16 == 0 && q.lda % 4 == 0 && q.ldb % 4 == 0 && (P.a_kmajor || P.M % 4 == 0) &&
                      (P.b_kmajor || P.N % 4 == 0) && P.kbs == 0;
      CHECK(ok, "problem %d runs direct tiles it cannot", i);
    }
    if(P.tiles_m * P.tiles_n == 0 || P.kbs == 0) {
      const bool none = !packs.count(std::make_pair(i, 0)) && !packs.count(std::make_pair(i, 1));
      CHECK(none, "problem %d without work is packed", i);
      continue;
    }
    if(P.prepacked) {
      const int pa = packs[std::make_pair(i, 1)], pb = packs[std::make_pair(i, 0)];
      CHECK(pa == 1 && pb == 1, "problem %d packed %d / %d times (A / B)", i, pa, pb);
    } else {
      const bool none = !packs.count(std::make_pair(i, 0)) && !packs.count(std::make_pair(i, 1));
      CHECK(none, "in-kernel-packed problem %d also in a job", i);
      std::set<std::tuple<int, int, int>> items;
      for(const GItem& g : pl.items)
        if((g.type == ITEM_PACK_A || g.type == ITEM_PACK_B) && g.problem == i) items.insert({g.type, g.x, g.y});
      CHECK((int)items.size() == ((int)ceil_div(P.M, 32) + (int)ceil_div(P.N, 32)) * P.pack_slices,
          "problem %d: %zu pack items", i, items.size());
    }
  }
}

// The triangular schedules' tile order: a bijection onto the lower triangle.
static void check_lower_order() {
  const std::string ctx = "triangle order";
  for(int group : {1, 3, 8})
    for(int T = 1; T <= 70; ++T) {
      const int n = T * (T + 1) / 2;
      std::set<std::pair<int, int>> seen;
      for(int p = 0; p < n; ++p) {
        int ti = -1, tj = -1;
        grouped_lower_tile(p, T, group, ti, tj);
        CHECK(tj >= 0 && tj <= ti && ti < T, "T %d group %d: p %d -> (%d, %d) outside the triangle", T, group, p, ti, tj);
        CHECK(seen.insert({ti, tj}).second, "T %d group %d: (%d, %d) twice", T, group, ti, tj);
      }
      CHECK((int)seen.size() == n, "T %d group %d: %zu of %d tiles", T, group, seen.size(), n);
    }
}

// A triangle whose sparse last round is re-tiled (tri_tail): every small
// tile of the lower triangle is computed exactly once, by a big tile of the
// first launch or by a sub-tile of the second.
static void check_triangle_tail() {
  const std::string ctx = "triangle tail";
  int splits = 0;
  for(mfx_precision p : {MFX_PREC_FP32, MFX_PREC_SPLIT_BF16, MFX_PREC_BF16})
    for(int cus : {256, 304, 80})
      for(int64_t n : {1000, 2048, 4096, 5000, 8192, 8200, 12288}) {
        for(int c = 0; c < config_count(p); ++c) {
          int bm, bn, bk, w;
          if(!config_info(p, c, &bm, &bn, &bk, &w) || bm != bn) continue;
          const int T = (int)ceil_div(n, bm);
          const TriTail tt = tri_tail(p, c, T, cus);
          const int tiles = T * (T + 1) / 2;
          if(tt.small < 0) {
            CHECK(tt.head == tiles, "no split but head %d of %d tiles", tt.head, tiles);
            continue;
          }
          ++splits;
          int sm, sn, sk, sw;
          CHECK(config_info(p, tt.small, &sm, &sn, &sk, &sw) && sm == sn && sm * tt.ratio == bm,
              "small configuration %d for %d / %d", tt.small, bm, tt.ratio);
          CHECK(tt.head % cus == 0 && tt.head > 0 && tt.head < tiles, "head %d of %d tiles", tt.head, tiles);
          const int r = tt.ratio, Ts = (int)ceil_div(n, bm / r);
          std::map<std::pair<int, int>, int> cover;
          for(int q = 0; q < tt.head; ++q) {
            int bi, bj;
            grouped_lower_tile(q, T, 8, bi, bj);
            for(int a = 0; a < r; ++a)
              for(int b = 0; b < r; ++b)
                if(bj * r + b <= bi * r + a && bi * r + a < Ts) ++cover[{bi * r + a, bj * r + b}];
          }
          for(int q = 0; q < (tiles - tt.head) * r * r; ++q) {
            int ti, tj;
            if(triangle_tile(q, Ts, 8, tt.head, r, T, ti, tj)) ++cover[{ti, tj}];
          }
          int missing = 0, twice = 0;
          for(int i = 0; i < Ts; ++i)
            for(int j = 0; j <= i; ++j) {
              auto it = cover.find({i, j});
              if(it == cover.end()) ++missing;
              else if(it->second != 1) ++twice;
            }
          CHECK(missing == 0 && twice == 0 && (int)cover.size() == Ts * (Ts + 1) / 2,
              "prec %d cus %d n %lld config %d: %d small tiles missing, %d covered twice", (int)p, cus, (long long)n, c,
              missing, twice);
        }
      }
  CHECK(splits > 10, "only %d triangles split", splits);
}

int main() {
  check_lower_order();
  check_triangle_tail();
  std::vector<Case> cases;
  {
    // the batched workload: uneven problems (log-uniform dims, multiples of
    // 16) plus an 8-long chain of 2048^2
    Case c{"uneven+chain", {}, {}};
    std::mt19937_64 rng(1);
    std::uniform_real_distribution<double> u(std::log(64.0), std::log(2048.0));
    for(int i = 0; i < 512; ++i) {
      int64_t d[3];
      for(auto& x : d) x = std::max<int64_t>(64, std::min<int64_t>(2048, (int64_t)std::llround(std::exp(u(rng)) / 16) * 16));
      c.probs.push_back(prob(d[0], d[1], d[2], i));
    }
    add_chain(c, 8, 2048);
    cases.push_back(c);
  }
  {
    Case c{"uneven-only", {}, {}};
    std::mt19937_64 rng(7);
    std::uniform_int_distribution<int> u(1, 700);
    for(int i = 0; i < 200; ++i) c.probs.push_back(prob(u(rng), u(rng), u(rng), i, i % 2, (i / 2) % 2));
    cases.push_back(c);
  }
  {
    Case c{"2mm", {prob(512, 512, 512, 0), prob(512, 512, 512, 1)}, {{1, 0}}};
    cases.push_back(c);
    Case d{"3mm", {prob(300, 200, 100, 0), prob(200, 250, 90, 1), prob(300, 250, 200, 2)}, {{2, 0}, {2, 1}}};
    cases.push_back(d);
  }
  {
    Case c{"chain5", {}, {}};
    add_chain(c, 5, 384);
    cases.push_back(c);
  }
  {
    Case c{"edges", {prob(64, 64, 0, 0), prob(0, 64, 64, 1), prob(64, 0, 64, 2), prob(1, 1, 1, 3), prob(33, 65, 17, 4),
                     prob(64, 64, 0, 5)},
        {{4, 3}, {5, 4}}};
    cases.push_back(c);
    Case e{"empty", {}, {}};
    cases.push_back(e);
    Case s{"single", {prob(4096, 64, 8192, 0)}, {}};
    cases.push_back(s);
  }
  int plans = 0;
  for(const Case& c : cases)
    for(mfx_precision prec : {MFX_PREC_FP32, MFX_PREC_SPLIT_BF16, MFX_PREC_BF16})
      for(int p = -1; p < grouped_policy_count(); ++p) {
        check_plan(c, prec, p);
        ++plans;
        if(g_failures > 20) return 1;
      }
  if(g_failures) {
    std::printf("FAILED %d of %d plans\n", g_failures, plans);
    return 1;
  }
  std::printf("OK %d\n", plans);
  return 0;
}
