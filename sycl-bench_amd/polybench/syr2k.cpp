// Polybench_Syr2k: C = BETA * C + ALPHA * (A * B^T + B * A^T)
// (polybench/syr2k.cpp:70-89), lower-triangle tiles with the mirror write, as
// for syrk; both products accumulate into the same tile.
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"

// polybench/syr2k.cpp:15-16
constexpr float ALPHA = 1;
constexpr float BETA = 1;

class Polybench_Syr2k {
public:
  explicit Polybench_Syr2k(const BenchmarkArgs& args)
      : args_(args), size_(args.problem_size), opt_(args),
        alpha_(args.cli.getOrDefault<float>("--alpha", opt_.reference_inputs() ? ALPHA : 1.5f)),
        beta_(args.cli.getOrDefault<float>("--beta", opt_.reference_inputs() ? BETA : 1.2f)) {}

  void setup() {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    if(opt_.reference_inputs()) {
      // polybench/syr2k.cpp:18-32
      initFromHost(d, C_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * j + 2) / n; });
      initFromHost(d, A_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * j) / n; });
      initFromHost(d, B_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * j + 1) / n; });
    } else {
      fillSeeded(opt_, d, A_, n, n, 0);
      fillSeeded(opt_, d, B_, n, n, 1);
      fillSeeded(opt_, d, C_, n, n, 2);
    }
    C0_.copyFrom(d, C_);
    mfxCheck(mfx_reserve_workspace(d.handle(), mfx_gemm_workspace_size(opt_.precision, n, n, n)), "workspace");
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    const int64_t n = size_;
    timedSubmit(d, kernels, [&] {
      mfxCheck(mfx_syr2k(d.handle(), opt_.precision, MFX_LOWER, 1, n, n, alpha_, A_.get(), n, B_.get(), n, beta_,
                   C_.get(), n, d.stream()),
          "mfx_syr2k");
    });
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    auto A = toF64(d, A_, n, n), B = toF64(d, B_, n, n);
    auto R = toF64(d, C0_, n, n);
    gemmF64(d, n, alpha_, A, false, B, true, beta_, R);
    gemmF64(d, n, alpha_, B, false, A, true, 1.0, R);
    check_ = checkAgainstF64(opt_, d, C_, R, n, n);
    return check_.pass;
  }

  void extraResults(std::map<std::string, std::string>& out) { contractionExtras(out, opt_, check_); }

  // Work of the triangle schedule: 2N(N+1)M (the reference's square does 4N^2 M).
  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const double n = static_cast<double>(args.problem_size);
    return {2.0 * n * (n + 1.0) * n, "FLOP"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const { return "Polybench_Syr2k" + opt_.suffix(); }

private:
  BenchmarkArgs args_;
  std::size_t size_;
  ContractionOptions opt_;
  float alpha_, beta_;
  DeviceBuffer<float> A_, B_, C_, C0_;
  CheckResult check_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<Polybench_Syr2k>();
  return app.exitCode();
}
