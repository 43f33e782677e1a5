// Polybench_3mm: E = A * B, F = C * D, G = E * F (polybench/3mm.cpp:112-159).
// Three mfx_gemm calls in stream order; the first two are independent, the
// third reads both (the reference's diamond DAG).
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"

class Polybench_3mm {
public:
  explicit Polybench_3mm(const BenchmarkArgs& args) : args_(args), size_(args.problem_size), opt_(args) {}

  void setup() {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    if(opt_.reference_inputs()) {
      // polybench/3mm.cpp:17-47
      initFromHost(d, A_, n, n, [n](std::size_t i, std::size_t k) { return ((float)i * k) / n; });
      initFromHost(d, B_, n, n, [n](std::size_t k, std::size_t j) { return ((float)k * (j + 1)) / n; });
      initFromHost(d, C_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * (j + 3)) / n; });
      initFromHost(d, D_, n, n, [n](std::size_t i, std::size_t j) { return ((float)i * (j + 2)) / n; });
    } else {
      fillSeeded(opt_, d, A_, n, n, 0);
      fillSeeded(opt_, d, B_, n, n, 1);
      fillSeeded(opt_, d, C_, n, n, 2);
      fillSeeded(opt_, d, D_, n, n, 3);
    }
    E_.initialize(d, n * n);
    F_.initialize(d, n * n);
    G_.initialize(d, n * n);
    mfxCheck(mfx_reserve_workspace(d.handle(), mfx_gemm_workspace_size(opt_.precision, n, n, n)), "workspace");
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    const int64_t n = size_;
    auto gemm = [&](const float* A, const float* B, float* C) {
      timedSubmit(d, kernels, [&] {
        mfxCheck(mfx_gemm(d.handle(), opt_.precision, MFX_OP_N, MFX_OP_N, n, n, n, 1.0f, A, n, B, n, 0.0f, C, n,
                     d.stream()),
            "mfx_gemm");
      });
    };
    gemm(A_.get(), B_.get(), E_.get());  // Polybench_3mm_1
    gemm(C_.get(), D_.get(), F_.get());  // Polybench_3mm_2
    gemm(E_.get(), F_.get(), G_.get());  // Polybench_3mm_3
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    auto A = toF64(d, A_, n, n), B = toF64(d, B_, n, n), C = toF64(d, C_, n, n), D = toF64(d, D_, n, n);
    DeviceBuffer<double> E, F, G;
    E.initialize(d, n * n);
    F.initialize(d, n * n);
    G.initialize(d, n * n);
    gemmF64(d, n, 1.0, A, false, B, false, 0.0, E);
    gemmF64(d, n, 1.0, C, false, D, false, 0.0, F);
    gemmF64(d, n, 1.0, E, false, F, false, 0.0, G);
    check_ = checkAgainstF64(opt_, d, G_, G, n, n, 2);
    return check_.pass;
  }

  void extraResults(std::map<std::string, std::string>& out) { contractionExtras(out, opt_, check_); }

  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const double n = static_cast<double>(args.problem_size);
    return {6.0 * n * n * n, "FLOP"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const { return "Polybench_3mm" + opt_.suffix(); }

private:
  BenchmarkArgs args_;
  std::size_t size_;
  ContractionOptions opt_;
  DeviceBuffer<float> A_, B_, C_, D_, E_, F_, G_;
  CheckResult check_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<Polybench_3mm>();
  return app.exitCode();
}
