// MicroBench_MFMA_<precision>: the matrix-core ceiling the contraction
// kernels are measured against.  The methodology of micro/arith.cpp:11-68 (a
// register-only loop whose result is known exactly, FLOP / time, verify every
// output) applied to the MFMA instructions the kernels use: mfx_mfma_burn runs
// two waves per SIMD on every CU, each issuing --size iterations of eight
// independent 32x32 MFMAs.
//   --precision=fp32 | bf16 | fp8-block   (split-bf16 runs the bf16 loop,
//                                          fp8-block2 the fp8 one)
//   --operands=constant | random          constant: power-of-two operands,
//       --size MFMAs per accumulator (upper bound); random: per-lane random
//       operands, 4 x --size + 1 MFMAs per accumulator (the power draw, and
//       clock, of a real contraction); name suffix "_random"
#include <string>
#include <vector>

#include "common.h"
#include "contraction.h"

class MicroBench_MFMA {
public:
  explicit MicroBench_MFMA(const BenchmarkArgs& args) : args_(args), opt_(args), random_(randomOperands(args)) {}

  static bool randomOperands(const BenchmarkArgs& args) {
    const std::string o = args.cli.getOrDefault<std::string>("--operands", "constant");
    if(o != "constant" && o != "random") throw std::invalid_argument("--operands must be constant or random");
    return o == "random";
  }

  void setup() {
    HipDevice& d = hipDevice(args_);
    waves_ = static_cast<std::size_t>(mfx_mfma_burn_waves(d.handle()));
    out_.initialize(d, waves_);
  }

  void run(std::vector<KernelTiming>& kernels) {
    HipDevice& d = hipDevice(args_);
    timedSubmit(d, kernels, [&] {
      mfxCheck(mfx_mfma_burn(d.handle(), opt_.precision, static_cast<int64_t>(args_.problem_size), random_ ? 1 : 0,
                   out_.get(),
                   static_cast<int64_t>(waves_), &flop_, &expected_, d.stream()),
          "mfx_mfma_burn");
    });
  }

  // Every wave's checksum is an exact power-of-two multiple.
  bool verify(VerificationSetting&) {
    HipDevice& d = hipDevice(args_);
    for(float v : out_.toHost(d))
      if(v != expected_) return false;
    return true;
  }

  static ThroughputMetric getThroughputMetric(const BenchmarkArgs& args) {
    const ContractionOptions o(args);
    const double per_mfma = o.precision == MFX_PREC_FP32 ? 2.0 * 32 * 32 * 2
                            : (o.precision == MFX_PREC_FP8_BLOCK || o.precision == MFX_PREC_FP8_BLOCK2)
                                ? 2.0 * 32 * 32 * 64
                                : 2.0 * 32 * 32 * 16;
    const double waves = static_cast<double>(mfx_mfma_burn_waves(hipDevice(args).handle()));
    const double n = static_cast<double>(args.problem_size);
    const double mfmas = randomOperands(args) ? 4.0 * n + 1.0 : n;
    return {waves * 8.0 * mfmas * per_mfma, "FLOP"};
  }

  std::string getBenchmarkName(const BenchmarkArgs&) const {
    std::string p = opt_.precision_name;
    if(opt_.precision == MFX_PREC_SPLIT_BF16) p = "bf16";
    if(opt_.precision == MFX_PREC_FP8_BLOCK2) p = "fp8-block";
    p.erase(std::remove(p.begin(), p.end(), '-'), p.end());
    return "MicroBench_MFMA_" + p + (random_ ? "_random" : "");
  }

private:
  BenchmarkArgs args_;
  ContractionOptions opt_;
  bool random_;
  std::size_t waves_ = 0;
  DeviceBuffer<float> out_;
  double flop_ = 0.0;
  float expected_ = 0.0f;
};

int main(int argc, char** argv) {
  BenchmarkApp app(argc, argv);
  app.run<MicroBench_MFMA>();
  return app.exitCode();
}
