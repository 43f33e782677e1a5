// Polybench_3mm: E = A * B, F = C * D, G = E * F (polybench/3mm.cpp:112-159).
// Three mfx_gemm calls in stream order; the first two are independent, the
// third reads both (the reference's diamond DAG).  --grouped submits the three
// as one persistent grouped launch with the DAG as in-kernel dependencies.
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"

class Polybench_3mm {
public:
  explicit Polybench_3mm(const BenchmarkArgs& args)
      : args_(args), size_(args.problem_size), opt_(args), grouped_(args.cli.isFlagSet("--grouped")) {}

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
    if(grouped_) {
      const mfx_gemm_problem p[3] = {
          {n, n, n, 1.0f, 0.0f, A_.get(), n, 0, B_.get(), n, 0, E_.get(), n},   // Polybench_3mm_1
          {n, n, n, 1.0f, 0.0f, C_.get(), n, 0, D_.get(), n, 0, F_.get(), n},   // Polybench_3mm_2
          {n, n, n, 1.0f, 0.0f, E_.get(), n, 0, F_.get(), n, 0, G_.get(), n}};  // Polybench_3mm_3
      const mfx_dep deps[2] = {{2, 0}, {2, 1}};
      timedSubmit(d, kernels, [&] {
        mfxCheck(mfx_gemm_grouped(d.handle(), opt_.precision, p, 3, deps, 2, d.stream()), "mfx_gemm_grouped");
      });
      return;
    }
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

  std::string getBenchmarkName(const BenchmarkArgs&) const {
    return "Polybench_3mm" + opt_.suffix() + (grouped_ ? "_grouped" : "");
  }

private:
  BenchmarkArgs args_;
  std::size_t size_;
  ContractionOptions opt_;
  bool grouped_;
  DeviceBuffer<float> A_, B_, C_, D_, E_, F_, G_;
  CheckResult check_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<Polybench_3mm>();
  return app.exitCode();
}
