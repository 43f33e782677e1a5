// Polybench_2mm.  With the reference's inputs (--init=polybench) it computes
// the reference's PolyBench/GPU form, polybench/2mm.cpp:91-124:
//   C += A * B;  E = C * D          (two contractions, the second reads C)
// With seeded inputs it computes the PolyBench 4 form used for the precision
// study, D = alpha * (A * B) * C + beta * D, as T = A * B; D = alpha*T*C + beta*D.
// Each contraction is one mfx_gemm call on the device stream, so the second
// waits for the first by stream order (the reference's accessor DAG edge);
// --grouped submits both as one persistent grouped launch in which the second
// waits for the first in-kernel.
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"

class Polybench_2mm {
public:
  explicit Polybench_2mm(const BenchmarkArgs& args)
      : args_(args), size_(args.problem_size), opt_(args), alpha_(args.cli.getOrDefault<float>("--alpha", 1.5f)),
        beta_(args.cli.getOrDefault<float>("--beta", 1.2f)), grouped_(args.cli.isFlagSet("--grouped")) {}

  void setup() {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    if(opt_.reference_inputs()) {
      // polybench/2mm.cpp:16-45
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
    C0_.copyFrom(d, C_);
    D0_.copyFrom(d, D_);
    mfxCheck(mfx_reserve_workspace(d.handle(), mfx_gemm_workspace_size(opt_.precision, n, n, n)), "workspace");
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    const int64_t n = size_;
    auto gemm = [&](float alpha, const float* A, const float* B, float beta, float* C) {
      timedSubmit(d, kernels, [&] {
        mfxCheck(mfx_gemm(d.handle(), opt_.precision, MFX_OP_N, MFX_OP_N, n, n, n, alpha, A, n, B, n, beta, C, n,
                     d.stream()),
            "mfx_gemm");
      });
    };
    if(grouped_) {
      mfx_gemm_problem p[2];
      if(opt_.reference_inputs()) {
        p[0] = {n, n, n, 1.0f, 1.0f, A_.get(), n, 0, B_.get(), n, 0, C_.get(), n};  // C += A B
        p[1] = {n, n, n, 1.0f, 0.0f, C_.get(), n, 0, D_.get(), n, 0, E_.get(), n};  // E = C D
      } else {
        p[0] = {n, n, n, 1.0f, 0.0f, A_.get(), n, 0, B_.get(), n, 0, E_.get(), n};      // T = A B
        p[1] = {n, n, n, alpha_, beta_, E_.get(), n, 0, C_.get(), n, 0, D_.get(), n};  // D = a T C + b D
      }
      const mfx_dep dep{1, 0};
      timedSubmit(d, kernels, [&] {
        mfxCheck(mfx_gemm_grouped(d.handle(), opt_.precision, p, 2, &dep, 1, d.stream()), "mfx_gemm_grouped");
      });
      return;
    }
    if(opt_.reference_inputs()) {
      gemm(1.0f, A_.get(), B_.get(), 1.0f, C_.get());  // Polybench_2mm_1, polybench/2mm.cpp:94-107
      gemm(1.0f, C_.get(), D_.get(), 0.0f, E_.get());  // Polybench_2mm_2, polybench/2mm.cpp:109-123
    } else {
      gemm(1.0f, A_.get(), B_.get(), 0.0f, E_.get());      // T = A * B (in E)
      gemm(alpha_, E_.get(), C_.get(), beta_, D_.get());   // D = alpha * T * C + beta * D
    }
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_;
    auto A = toF64(d, A_, n, n), B = toF64(d, B_, n, n), C = toF64(d, C0_, n, n), D = toF64(d, D0_, n, n);
    if(opt_.reference_inputs()) {
      gemmF64(d, n, 1.0, A, false, B, false, 1.0, C);  // C += A B
      DeviceBuffer<double> E;
      E.initialize(d, n * n);
      gemmF64(d, n, 1.0, C, false, D, false, 0.0, E);  // E = C D
      check_ = checkAgainstF64(opt_, d, E_, E, n, n, 2);
    } else {
      DeviceBuffer<double> T;
      T.initialize(d, n * n);
      gemmF64(d, n, 1.0, A, false, B, false, 0.0, T);
      gemmF64(d, n, alpha_, T, false, C, false, beta_, D);
      check_ = checkAgainstF64(opt_, d, D_, D, n, n, 2);
    }
    return check_.pass;
  }

  void extraResults(std::map<std::string, std::string>& out) { contractionExtras(out, opt_, check_); }

  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const double n = static_cast<double>(args.problem_size);
    return {4.0 * n * n * n, "FLOP"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const {
    return "Polybench_2mm" + opt_.suffix() + (grouped_ ? "_grouped" : "");
  }

private:
  BenchmarkArgs args_;
  std::size_t size_;
  ContractionOptions opt_;
  float alpha_, beta_;
  bool grouped_;
  DeviceBuffer<float> A_, B_, C_, D_, E_, C0_, D0_;
  CheckResult check_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<Polybench_2mm>();
  return app.exitCode();
}
