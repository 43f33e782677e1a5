// Polybench_Syrk: C = beta * C + alpha * A * A^T (polybench/syrk.cpp:67-85).
// The reference computes the full N x N square; here only lower-triangle
// tiles are scheduled (mfx_syrk, MFX_LOWER) and each strictly-lower element
// is also written to its mirror position from that position's own C value,
// so the full square -- what the reference's verify() reads -- is produced
// with half the contraction work.
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"

// polybench/syrk.cpp:15-16
constexpr float alpha = 123;
constexpr float beta = 14512;

class Polybench_Syrk {
public:
  explicit Polybench_Syrk(const BenchmarkArgs& args)
      : args_(args), size_(args.problem_size), opt_(args),
        alpha_(args.cli.getOrDefault<float>("--alpha", opt_.reference_inputs() ? alpha : 1.5f)),
        beta_(args.cli.getOrDefault<float>("--beta", opt_.reference_inputs() ? beta : 1.2f)) {}

  void setup() {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    if(opt_.reference_inputs()) {
      // polybench/syrk.cpp:18-31
      initFromHost(d, A_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * j) / n; });
      initFromHost(d, C_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * j + 2) / n; });
    } else {
      fillSeeded(opt_, d, A_, n, n, 0);
      fillSeeded(opt_, d, C_, n, n, 2);
    }
    C0_.copyFrom(d, C_);
    mfxCheck(mfx_reserve_workspace(d.handle(), mfx_gemm_workspace_size(opt_.precision, n, n, n)), "workspace");
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    const int64_t n = size_;
    timedSubmit(d, kernels, [&] {
      mfxCheck(mfx_syrk(d.handle(), opt_.precision, MFX_LOWER, 1, n, n, alpha_, A_.get(), n, beta_, C_.get(), n,
                   d.stream()),
          "mfx_syrk");
    });
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    auto A = toF64(d, A_, n, n);
    auto R = toF64(d, C0_, n, n);
    gemmF64(d, n, alpha_, A, false, A, true, beta_, R);
    check_ = checkAgainstF64(opt_, d, C_, R, n, n);
    return check_.pass;
  }

  void extraResults(std::map<std::string, std::string>& out) { contractionExtras(out, opt_, check_); }

  // Work of the triangle schedule: N(N+1)M (the reference's square does 2N^2 M).
  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const double n = static_cast<double>(args.problem_size);
    return {n * (n + 1.0) * n, "FLOP"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const { return "Polybench_Syrk" + opt_.suffix(); }

private:
  BenchmarkArgs args_;
  std::size_t size_;
  ContractionOptions opt_;
  float alpha_, beta_;
  DeviceBuffer<float> A_, C_, C0_;
  CheckResult check_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<Polybench_Syrk>();
  return app.exitCode();
}
