// Polybench_Correlation: column means and standard deviations, centring and
// scaling, the upper-triangle correlation matrix symmat = dataᵀ·data,
// mirrored, with a unit diagonal (polybench/correlation.cpp:92-158), as one
// mfx_correlation call: three HBM passes, then the N x N x N product on the
// triangle contraction kernel (with --precision, split / single BF16).
//
// Storage follows the reference: (N+1) x (N+1) arrays used 1-based (ld =
// N+1).  Its initialisation writes the flat array with a row stride of N, not
// N+1 (correlation.cpp:17-25), which this program reproduces, so the values
// at (i, j) are those of the reference.  Deliberate differences: mean and
// stddev are computed afresh and the centred data goes to a work array, so
// repeated runs agree (the reference accumulates onto the previous run's mean
// and stddev and centres in place).
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"
#include "polybenchUtilFuncts.h"
#include "statistics.h"

// polybench/correlation.cpp:8-9
constexpr double FLOAT_N = 3214212.01;
constexpr double EPS = 0.005;

class Polybench_Correlation {
public:
  explicit Polybench_Correlation(const BenchmarkArgs& args) : args_(args), size_(args.problem_size), opt_(args) {}

  void setup() {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_, n1 = size_ + 1;
    if(opt_.reference_inputs()) {
      // polybench/correlation.cpp:17-25 (flat index i * N + j, i, j = 0..N)
      host_.assign(n1 * n1, 0.0f);
      for(std::size_t i = 0; i <= n; i++)
        for(std::size_t j = 0; j <= n; j++) host_[i * n + j] = ((float)i * j) / (n + 1);
      data_.initialize(d, host_.data(), host_.size());
    } else {
      data_.initialize(d, n1 * n1);
      mfxCheck(mfx_fill_uniform(d.handle(), n, n, interior(data_, n1), n1, opt_.seed, 0, d.stream()), "fill");
    }
    mean_.initialize(d, n1);
    stddev_.initialize(d, n1);
    centered_.initialize(d, n1 * n1);
    symmat_.initialize(d, n1 * n1);
    mfxCheck(mfx_reserve_workspace(d.handle(), mfx_gemm_workspace_size(opt_.precision, n, n, n)), "workspace");
    d.synchronize();
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    const int64_t n = size_, n1 = size_ + 1;
    timedSubmit(d, kernels, [&] {
      mfxCheck(mfx_correlation(d.handle(), opt_.precision, n, n, interior(data_, n1), n1, FLOAT_N, EPS,
                   mean_.get() + 1, stddev_.get() + 1, interior(centered_, n1), n1, interior(symmat_, n1), n1,
                   d.stream()),
          "mfx_correlation");
    });
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    const std::size_t n = size_, n1 = size_ + 1;
    const std::vector<float> x = hostInterior(d, data_, n1, n, n);
    auto R = statsF64(d, x, n, n, n, FLOAT_N, true, EPS);
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

  std::string getBenchmarkName(const BenchmarkArgs&) const { return "Polybench_Correlation" + opt_.suffix(); }

private:
  BenchmarkArgs args_;
  std::size_t size_;
  ContractionOptions opt_;
  std::vector<float> host_;
  DeviceBuffer<float> data_, mean_, stddev_, centered_, symmat_;
  CheckResult check_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<Polybench_Correlation>();
  return app.exitCode();
}
