// Polybench_Covariance: column means, centring and the upper-triangle
// covariance matrix symmat = dataᵀ·data, mirrored (polybench/covariance.cpp:
// 76-128), as one mfx_covariance call: two HBM passes for the means, one for
// centring, and the N x N x N product on the triangle contraction kernel
// (with --precision, split / single BF16).
//
// Storage follows the reference: (N+1) x (N+1) arrays used 1-based, so the
// computed block is rows / columns 1..N (ld = N+1).  Deliberate difference:
// the centred data goes to a work array, so `data` keeps its inputs and
// repeated runs give the same result (the reference centres in place).
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"
#include "statistics.h"

// polybench/covariance.cpp:18
constexpr float float_n = 3214212.01f;

class Polybench_Covariance {
public:
  explicit Polybench_Covariance(const BenchmarkArgs& args) : args_(args), size_(args.problem_size), opt_(args) {}

  void setup() {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_, n1 = size_ + 1;
    if(opt_.reference_inputs()) {
      // polybench/covariance.cpp:20-29: rows 0..N-1 and columns 0..N-1 of the
      // (N+1)-wide array, the rest zero
      host_.assign(n1 * n1, 0.0f);
      for(std::size_t i = 0; i < n; i++)
        for(std::size_t j = 0; j < n; j++) host_[i * n1 + j] = ((float)i * j) / n;
      data_.initialize(d, host_.data(), host_.size());
    } else {
      data_.initialize(d, n1 * n1);
      mfxCheck(mfx_fill_uniform(d.handle(), n, n, interior(data_, n1), n1, opt_.seed, 0, d.stream()), "fill");
    }
    mean_.initialize(d, n1);
    centered_.initialize(d, n1 * n1);
    symmat_.initialize(d, n1 * n1);
    mfxCheck(mfx_reserve_workspace(d.handle(), mfx_gemm_workspace_size(opt_.precision, n, n, n)), "workspace");
    d.synchronize();
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    const int64_t n = size_, n1 = size_ + 1;
    timedSubmit(d, kernels, [&] {
      mfxCheck(mfx_covariance(d.handle(), opt_.precision, n, n, interior(data_, n1), n1, float_n, mean_.get() + 1,
                   interior(centered_, n1), n1, interior(symmat_, n1), n1, d.stream()),
          "mfx_covariance");
    });
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_, n1 = size_ + 1;
    const std::vector<float> x = hostInterior(d, data_, n1, n, n);
    auto R = statsF64(d, x, n, n, n, float_n, false, 0.0);
    auto X = denseInterior(d, symmat_, n1, n);
    check_ = checkAgainstF64(opt_, d, X, R, n, n);
    return check_.pass;
  }

  void extraResults(std::map<std::string, std::string>& out) { contractionExtras(out, opt_, check_); }

  // Upper-triangle product: N(N+1)/2 dot products of length N.
  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const double n = static_cast<double>(args.problem_size);
    return {n * (n + 1.0) * n, "FLOP"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const { return "Polybench_Covariance" + opt_.suffix(); }

private:
  BenchmarkArgs args_;
  std::size_t size_;
  ContractionOptions opt_;
  std::vector<float> host_;
  DeviceBuffer<float> data_, mean_, centered_, symmat_;
  CheckResult check_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<Polybench_Covariance>();
  return app.exitCode();
}
