// Polybench_Gemm: C = alpha * A * B + beta * C on N x N FP32 matrices, one
// mfx_gemm call per run.  Mirrors polybench/gemm.cpp of the reference: the
// same class name, constructor, setup()/run()/verify()/getBenchmarkName(),
// the same inputs and constants by default (--init=polybench), and the same
// verify rule (percentDiff > 0.05 fails); the expected values come from an
// FP64 contraction on the GPU instead of the reference's serial CPU loop.
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"

// polybench/gemm.cpp:11-12 (integer macros, promoted to float in use)
#define ALPHA 32412
#define BETA 2123

class Polybench_Gemm {
public:
  explicit Polybench_Gemm(const BenchmarkArgs& args)
      : args_(args), size_(args.problem_size), opt_(args),
        alpha_(args.cli.getOrDefault<float>("--alpha", opt_.reference_inputs() ? ALPHA : 1.5f)),
        beta_(args.cli.getOrDefault<float>("--beta", opt_.reference_inputs() ? BETA : 1.2f)) {}

  void setup() {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    if(opt_.reference_inputs()) {
      // polybench/gemm.cpp:18-40
      initFromHost(d, A_, n, n, [n](std::size_t i, std::size_t k) { return ((float)i * k) / n; });
      initFromHost(d, B_, n, n, [n](std::size_t k, std::size_t j) { return ((float)k * j + 1) / n; });
      initFromHost(d, C_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * j + 2) / n; });
    } else {
      fillSeeded(opt_, d, A_, n, n, 0);
      fillSeeded(opt_, d, B_, n, n, 1);
      fillSeeded(opt_, d, C_, n, n, 2);
    }
    C0_.copyFrom(d, C_);  // input C, kept for verify()
    mfxCheck(mfx_reserve_workspace(d.handle(), mfx_gemm_workspace_size(opt_.precision, n, n, n)), "workspace");
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    const int64_t n = size_;
    timedSubmit(d, kernels, [&] {
      mfxCheck(mfx_gemm(d.handle(), opt_.precision, MFX_OP_N, MFX_OP_N, n, n, n, alpha_, A_.get(), n, B_.get(), n,
                   beta_, C_.get(), n, d.stream()),
          "mfx_gemm");
    });
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    auto A = toF64(d, A_, n, n), B = toF64(d, B_, n, n);
    auto R = toF64(d, C0_, n, n);
    gemmF64(d, n, alpha_, A, false, B, false, beta_, R);
    check_ = checkAgainstF64(opt_, d, C_, R, n, n);
    return check_.pass;
  }

  void extraResults(std::map<std::string, std::string>& out) { contractionExtras(out, opt_, check_); }

  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const double n = static_cast<double>(args.problem_size);
    return {2.0 * n * n * n, "FLOP"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const { return "Polybench_Gemm" + opt_.suffix(); }

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
  app.run<Polybench_Gemm>();
  return app.exitCode();
}
