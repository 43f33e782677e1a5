// The autotuner: times every compiled tile configuration of a precision on a
// problem shape and records the fastest per (precision, shape bucket); for
// grouped launches it times every tile policy on the caller's own problem
// list and records the fastest for that launch's key.  Choices live in a process-wide table (mutex-protected) that
// mfx_gemm / mfx_syrk / mfx_syr2k / grouped launches consult, and can be
// saved to and loaded from a JSON file.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "mfx_internal.h"

namespace mfx {

std::string tune_key(mfx_precision p, int64_t m, int64_t n, int64_t k);
void record_tuned(const std::string& key, int config);
std::map<std::string, int> tuned_snapshot();
void replace_tuned(const std::map<std::string, int>& t);
mfx_status dispatch_dense(Handle* h, mfx_precision prec, DenseOp& op, int config, hipStream_t s);
int grouped_policy_count();
bool grouped_policy_info(int id, int* big, int* small, double* bound, int* per_level);
bool grouped_policy_direct(int id);
std::string grouped_tune_key(mfx_precision p, const mfx_gemm_problem* problems, int count, const mfx_dep* deps,
    int ndeps);
mfx_status time_grouped(Handle* h, mfx_precision prec, const mfx_gemm_problem* problems, int count,
    const mfx_dep* deps, int ndeps, int reps, hipStream_t s, double* ms);
mfx_status launch_fill_uniform(Handle* h, int64_t rows, int64_t cols, float* p, int64_t ld, uint64_t seed,
    uint32_t mid, hipStream_t s);

namespace {

// RAII device buffer for the tuner's scratch operands.
struct Scratch {
  void* p = nullptr;
  ~Scratch() {
    if(p) (void)hipFree(p);
  }
};

// Average milliseconds of `reps` calls of `run` after one warm-up call.
template <class F>
mfx_status time_ms(hipStream_t s, int reps, F&& run, double* ms) {
  MFX_CHECK(run());
  hipEvent_t a, b;
  MFX_HIP(hipEventCreate(&a));
  MFX_HIP(hipEventCreate(&b));
  mfx_status st = MFX_SUCCESS;
  if(hipEventRecord(a, s) != hipSuccess) st = MFX_ERROR_HIP;
  for(int r = 0; r < reps && st == MFX_SUCCESS; ++r) st = run();
  float t = 0.0f;
  if(st == MFX_SUCCESS) {
    hipError_t e = hipEventRecord(b, s);
    if(e == hipSuccess) e = hipEventSynchronize(b);
    if(e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if(e != hipSuccess) st = hip_fail(e, "autotune timing");
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = t / reps;
  return st;
}

std::string json_escape(const std::string& s) {
  std::string o;
  for(char c : s) {
    if(c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o;
}

}  // namespace

}  // namespace mfx

using namespace mfx;

extern "C" {

mfx_status mfx_autotune(mfx_handle handle, mfx_precision prec, int64_t m, int64_t n, int64_t k, int32_t* best_config,
    mfx_stream stream) {
  if(!handle) return invalid("null handle");
  if(!best_config) return invalid("null output pointer");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  if(prec < MFX_PREC_FP32 || prec > MFX_PREC_FP8_BLOCK2) return invalid("unknown precision");
  if(m <= 0 || n <= 0 || k <= 0 || m > (1 << 20) || n > (1 << 20) || k > (1 << 20))
    return invalid("autotune needs positive dimensions up to 2^20");
  const int count = config_count(prec);
  if(count == 0) return unsupported("no compiled configuration for this precision");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Scratch A, B, C;
  MFX_HIP(hipMalloc(&A.p, (size_t)m * k * 4));
  MFX_HIP(hipMalloc(&B.p, (size_t)k * n * 4));
  MFX_HIP(hipMalloc(&C.p, (size_t)m * n * 4));
  MFX_CHECK(launch_fill_uniform(h, m, k, static_cast<float*>(A.p), k, 7, 0, s));
  MFX_CHECK(launch_fill_uniform(h, k, n, static_cast<float*>(B.p), n, 7, 1, s));
  // Enough repetitions for ~20 GFLOP per configuration, 1 to 10.
  const double flop = 2.0 * m * n * k;
  const int reps = (int)std::max(1.0, std::min(10.0, 2e10 / flop));
  const bool timing = h->timing;
  h->timing = false;
  int best = -1;
  double best_ms = 0.0;
  mfx_status st = MFX_SUCCESS;
  for(int c = 0; c < count && st == MFX_SUCCESS; ++c) {
    DenseOp op{};
    op.M = (int)m;
    op.N = (int)n;
    op.K = (int)k;
    op.A = static_cast<const float*>(A.p);
    op.lda = k;
    op.a_kmajor = true;
    op.B = static_cast<const float*>(B.p);
    op.ldb = n;
    op.b_kmajor = false;
    op.nprod = 1;
    op.tiles = TILES_FULL;
    op.ep = Epilogue{static_cast<float*>(C.p), n, 1.0f, 0.0f, (int)m, (int)n, TILES_FULL, 0};
    double ms = 0.0;
    st = time_ms(s, reps, [&] { return dispatch_dense(h, prec, op, c, s); }, &ms);
    if(st == MFX_SUCCESS && (best < 0 || ms < best_ms)) {
      best = c;
      best_ms = ms;
    }
  }
  h->timing = timing;
  MFX_CHECK(st);
  record_tuned(tune_key(prec, m, n, k), best);
  *best_config = best;
  return MFX_SUCCESS;
}

mfx_status mfx_autotune_grouped(mfx_handle handle, mfx_precision prec, const mfx_gemm_problem* problems,
    int32_t count, const mfx_dep* deps, int32_t ndeps, int32_t* best_policy, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  if(prec != MFX_PREC_FP32 && prec != MFX_PREC_SPLIT_BF16 && prec != MFX_PREC_BF16)
    return unsupported("grouped launch supports FP32, SPLIT_BF16 and BF16");
  if(h->forced_config[prec] >= 0) return invalid("a forced configuration overrides grouped tuning; clear it first");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool timing = h->timing;
  const int forced_policy = h->grouped_policy;
  h->timing = false;
  std::vector<int> cand;
  for(int p = 0; p < grouped_policy_count(); ++p) {
    // FP32 runs 256x128 for the 256x256 configuration: time each distinct
    // pair of configurations once.
    auto pair_of = [&](int q) {
      int big, small, per_level;
      double bound;
      grouped_policy_info(q, &big, &small, &bound, &per_level);
      if(prec == MFX_PREC_FP32) {
        big = big == 0 ? 3 : big;
        small = small == 0 ? 3 : small;
      }
      const bool direct = prec != MFX_PREC_FP32 && grouped_policy_direct(q);
      return std::make_tuple(big, small, big == small ? 0.0 : bound, per_level, direct);
    };
    bool dup = false;
    for(int q = 0; q < p; ++q) dup = dup || pair_of(q) == pair_of(p);
    // without dependencies the per-level form is the same single launch
    if(!dup && !(ndeps == 0 && std::get<3>(pair_of(p)) == 1)) cand.push_back(p);
  }
  // Each policy planned and bound once, then timed over back-to-back
  // launches of the same tables; two rounds over the candidates, each
  // policy's faster round kept (the first round's early candidates can run
  // while the clock is still ramping up).
  std::vector<double> cand_ms(cand.size(), 0.0);
  mfx_status st = MFX_SUCCESS;
  for(int round = 0; round < 2 && st == MFX_SUCCESS; ++round)
    for(size_t i = 0; i < cand.size() && st == MFX_SUCCESS; ++i) {
      h->grouped_policy = cand[i];
      double ms = 0.0;
      st = time_grouped(h, prec, problems, count, deps, ndeps, 5, s, &ms);
      if(st == MFX_SUCCESS && (round == 0 || ms < cand_ms[i])) cand_ms[i] = ms;
    }
  int best = -1;
  double best_ms = 0.0;
  for(size_t i = 0; i < cand.size(); ++i)
    if(best < 0 || cand_ms[i] < best_ms) {
      best = cand[i];
      best_ms = cand_ms[i];
    }
  h->grouped_policy = forced_policy;
  h->timing = timing;
  MFX_CHECK(st);
  record_tuned(grouped_tune_key(prec, problems, count, deps, ndeps), best);
  if(best_policy) *best_policy = best;
  return MFX_SUCCESS;
}

int32_t mfx_grouped_policy_count(void) { return grouped_policy_count(); }

mfx_status mfx_force_grouped_policy(mfx_handle handle, int32_t policy) {
  if(!handle) return invalid("null handle");
  if(policy < -1 || policy >= grouped_policy_count()) return invalid("grouped policy out of range");
  handle->h.grouped_policy = policy;
  return MFX_SUCCESS;
}

mfx_status mfx_grouped_policy_info(int32_t policy, int32_t* large_config, int32_t* small_config,
    double* waste_bound, int32_t* per_level) {
  if(!large_config || !small_config || !waste_bound || !per_level) return invalid("null output pointer");
  int big, small, lv;
  double bound;
  if(!grouped_policy_info(policy, &big, &small, &bound, &lv)) return invalid("grouped policy out of range");
  *large_config = big;
  *small_config = small;
  *waste_bound = bound;
  *per_level = lv;
  return MFX_SUCCESS;
}

mfx_status mfx_autotune_save(const char* path) {
  if(!path) return invalid("null path");
  std::ofstream f(path);
  if(!f) {
    set_last_error(std::string("cannot write ") + path);
    return MFX_ERROR_IO;
  }
  f << "{\n  \"format\": \"mfx-autotune-2\",\n  \"entries\": {";
  bool first = true;
  for(const auto& kv : tuned_snapshot()) {
    f << (first ? "\n" : ",\n") << "    \"" << json_escape(kv.first) << "\": " << kv.second;
    first = false;
  }
  f << "\n  }\n}\n";
  if(!f) {
    set_last_error(std::string("write failed: ") + path);
    return MFX_ERROR_IO;
  }
  return MFX_SUCCESS;
}

mfx_status mfx_autotune_load(const char* path) {
  if(!path) return invalid("null path");
  std::ifstream f(path);
  if(!f) {
    set_last_error(std::string("cannot read ") + path);
    return MFX_ERROR_IO;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  if(text.find("\"mfx-autotune-2\"") == std::string::npos) {
    set_last_error(std::string(path) + " is not an mfx autotune file");
    return MFX_ERROR_IO;
  }
  std::map<std::string, int> table;
  const size_t e = text.find("\"entries\"");
  const size_t open = e == std::string::npos ? e : text.find('{', e);
  const size_t close = open == std::string::npos ? open : text.find('}', open);
  if(close == std::string::npos) {
    set_last_error(std::string(path) + ": no entries object");
    return MFX_ERROR_IO;
  }
  size_t pos = open + 1;
  while(true) {
    const size_t q0 = text.find('"', pos);
    if(q0 == std::string::npos || q0 > close) break;
    const size_t q1 = text.find('"', q0 + 1);
    const size_t colon = text.find(':', q1);
    if(q1 == std::string::npos || colon == std::string::npos || colon > close) {
      set_last_error(std::string(path) + ": malformed entry");
      return MFX_ERROR_IO;
    }
    char* end = nullptr;
    const long v = std::strtol(text.c_str() + colon + 1, &end, 10);
    if(end == text.c_str() + colon + 1 || v < 0 || v > 64) {
      set_last_error(std::string(path) + ": bad configuration id");
      return MFX_ERROR_IO;
    }
    table[text.substr(q0 + 1, q1 - q0 - 1)] = (int)v;
    pos = end - text.c_str();
  }
  replace_tuned(table);
  return MFX_SUCCESS;
}

}  // extern "C"
