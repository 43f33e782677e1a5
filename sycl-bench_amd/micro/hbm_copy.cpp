// MicroBench_DRAM_<type>_1: the HBM ceiling the contraction kernels are
// measured against.  micro/DRAM.cpp:27-96 with one-dimensional buffers: copy
// --size^3 bytes of 33 from one device buffer to another, report
// 2 x the copied GiB (read + write), verify that every element is 33.  The
// copy is mfx_hbm_copy (128-bit loads and stores, one chunk per
// thread).  The reference's 2-D and 3-D variants copy the same contiguous
// bytes through a different index space and have no counterpart here.
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"

template <class T>
class MicroBenchHBM {
public:
  explicit MicroBenchHBM(const BenchmarkArgs& args) : args_(args), count_(elements(args)) {}

  static std::size_t elements(const BenchmarkArgs& args) {
    const std::size_t s = args.problem_size;
    return s * s * s / sizeof(T);
  }

  void setup() {
    HipDevice& d = hipDevice(args_);
    std::vector<T> host(count_, static_cast<T>(33));
    in_.initialize(d, host.data(), host.size());
    out_.initialize(d, count_);
    d.synchronize();
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    timedSubmit(d, kernels, [&] {
      mfxCheck(mfx_hbm_copy(d.handle(), in_.get(), out_.get(), static_cast<int64_t>(count_ * sizeof(T)), d.stream()),
          "mfx_hbm_copy");
    });
  }

  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    for(T v : out_.toHost(d))
      if(v != static_cast<T>(33)) return false;
    return true;
  }

  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const double gib = static_cast<double>(elements(args) * sizeof(T)) / 1024.0 / 1024.0 / 1024.0;
    return {gib * 2.0, "GiB"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const {
    return std::string("MicroBench_DRAM_") + (sizeof(T) == 4 ? "float" : "double") + "_1";
  }

private:
  BenchmarkArgs args_;
  std::size_t count_;
  DeviceBuffer<T> in_, out_;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<MicroBenchHBM<float>>();
  app.run<MicroBenchHBM<double>>();
  return app.exitCode();
}
