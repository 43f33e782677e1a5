// Shared pieces of the dense-contraction benchmarks: benchmark-local options,
// variant naming, input generation on the device, and verification against an
// FP64 result computed on the GPU (mfx_gemm_f64).
//
// Options (read with the reference's cli.getOrDefault pattern):
//   --precision=fp32|split-bf16|bf16|fp8-block|fp8-block2   (default fp32)
//   --init=polybench|uniform|lognormal   polybench = the reference's own
//                                        closed-form inputs and constants
//   --seed=<u64>  --alpha=<f>  --beta=<f>  --tolerance=<f>
//   --metrics=extended   adds precision, init, seed, error and percentDiff columns
// The reference names (Polybench_Gemm, ...) belong to fp32 + polybench; any
// other combination appends "_<precision>" and/or "_<init>" so one program
// can run several variants (duplicate names are rejected).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "command_line.h"
#include "hip_device.h"
#include "memory_wrappers.h"
#include "mfx.h"
#include "time_metrics.h"

struct ContractionOptions {
  mfx_precision precision = MFX_PREC_FP32;
  std::string precision_name = "fp32";
  std::string init = "polybench";
  uint64_t seed = 1;
  double tolerance = -1.0;  // < 0: the precision's default

  explicit ContractionOptions(const BenchmarkArgs& a) {
    precision_name = a.cli.getOrDefault<std::string>("--precision", "fp32");
    if(precision_name == "fp32") precision = MFX_PREC_FP32;
    else if(precision_name == "split-bf16") precision = MFX_PREC_SPLIT_BF16;
    else if(precision_name == "bf16") precision = MFX_PREC_BF16;
    else if(precision_name == "fp8-block") precision = MFX_PREC_FP8_BLOCK;
    else if(precision_name == "fp8-block2") precision = MFX_PREC_FP8_BLOCK2;
    else throw std::invalid_argument("--precision must be fp32, split-bf16, bf16, fp8-block or fp8-block2");
    init = a.cli.getOrDefault<std::string>("--init", "polybench");
    if(init != "polybench" && init != "uniform" && init != "lognormal")
      throw std::invalid_argument("--init must be polybench, uniform or lognormal");
    seed = a.cli.getOrDefault<uint64_t>("--seed", 1);
    tolerance = a.cli.getOrDefault<double>("--tolerance", -1.0);
  }

  bool reference_inputs() const { return init == "polybench"; }

  // Suffix appended to the reference benchmark name for non-default variants.
  std::string suffix() const {
    std::string s;
    if(precision != MFX_PREC_FP32) {
      std::string p = precision_name;
      p.erase(std::remove(p.begin(), p.end(), '-'), p.end());
      s += "_" + p;
    }
    if(init != "polybench") s += "_" + init;
    return s;
  }

  // Relative Frobenius error allowed against FP64 for `stages` chained products.
  double frobenius_tolerance(int stages = 1) const {
    if(tolerance >= 0.0) return tolerance;
    const double chain = std::sqrt(static_cast<double>(stages));
    switch(precision) {
      case MFX_PREC_FP32: return 1e-5 * chain;
      case MFX_PREC_SPLIT_BF16: return 1e-5 * chain;
      case MFX_PREC_BF16: return 2e-2 * chain;
      case MFX_PREC_FP8_BLOCK: return 0.25;
      case MFX_PREC_FP8_BLOCK2: return 0.05;
    }
    return 0.0;
  }

  // Whether the reference's element-wise percentDiff rule applies: the
  // reference's own inputs at a precision that carries FP32 accuracy.
  bool percent_diff_rule() const {
    return reference_inputs() && (precision == MFX_PREC_FP32 || precision == MFX_PREC_SPLIT_BF16);
  }
};

// Fills a rows x cols row-major device matrix from a host function f(i, j).
template <class F>
void initFromHost(HipDevice& d, DeviceBuffer<float>& buf, std::size_t rows, std::size_t cols, F&& f) {
  std::vector<float> h(rows * cols);
  for(std::size_t i = 0; i < rows; ++i)
    for(std::size_t j = 0; j < cols; ++j) h[i * cols + j] = f(i, j);
  buf.initialize(d, h.data(), h.size());
  d.synchronize();  // the host vector dies here
}

inline void fillUniform(HipDevice& d, DeviceBuffer<float>& buf, std::size_t rows, std::size_t cols, uint64_t seed,
    uint32_t matrix_id) {
  buf.initialize(d, rows * cols);
  mfxCheck(mfx_fill_uniform(d.handle(), rows, cols, buf.get(), cols, seed, matrix_id, d.stream()), "mfx_fill_uniform");
}

inline void fillLognormal(HipDevice& d, DeviceBuffer<float>& buf, std::size_t rows, std::size_t cols, uint64_t seed,
    uint32_t matrix_id) {
  buf.initialize(d, rows * cols);
  mfxCheck(mfx_fill_lognormal(d.handle(), rows, cols, buf.get(), cols, seed, matrix_id, 0.0f, 2.0f, d.stream()),
      "mfx_fill_lognormal");
}

// Seeded input per --init (uniform or lognormal).
inline void fillSeeded(const ContractionOptions& o, HipDevice& d, DeviceBuffer<float>& buf, std::size_t rows,
    std::size_t cols, uint32_t matrix_id) {
  if(o.init == "lognormal")
    fillLognormal(d, buf, rows, cols, o.seed, matrix_id);
  else
    fillUniform(d, buf, rows, cols, o.seed, matrix_id);
}

// rows x cols FP64 copy of a float device matrix.
inline DeviceBuffer<double> toF64(HipDevice& d, const DeviceBuffer<float>& src, std::size_t rows, std::size_t cols) {
  DeviceBuffer<double> out;
  out.initialize(d, rows * cols);
  mfxCheck(mfx_convert_f32_f64(d.handle(), rows, cols, src.get(), cols, out.get(), cols, d.stream()),
      "mfx_convert_f32_f64");
  return out;
}

// C = alpha * op(A) op(B) + beta * C in FP64 on the device (square n x n operands).
inline void gemmF64(HipDevice& d, std::size_t n, double alpha, const DeviceBuffer<double>& A, bool ta,
    const DeviceBuffer<double>& B, bool tb, double beta, DeviceBuffer<double>& C) {
  mfxCheck(mfx_gemm_f64(d.handle(), ta ? MFX_OP_T : MFX_OP_N, tb ? MFX_OP_T : MFX_OP_N, n, n, n, alpha, A.get(), n,
               B.get(), n, beta, C.get(), n, d.stream()),
      "mfx_gemm_f64");
}

// Outcome of comparing a float result against its FP64 reference.
struct CheckResult {
  double rel_frobenius = 0.0;
  int64_t percent_diff_failures = 0;
  bool pass = false;
};

inline CheckResult checkAgainstF64(const ContractionOptions& o, HipDevice& d, const DeviceBuffer<float>& X,
    const DeviceBuffer<double>& R, std::size_t rows, std::size_t cols, int stages = 1) {
  CheckResult c;
  mfxCheck(mfx_rel_frobenius(d.handle(), MFX_FULL, rows, cols, X.get(), cols, R.get(), cols, &c.rel_frobenius,
               d.stream()),
      "mfx_rel_frobenius");
  if(o.percent_diff_rule()) {
    // ERROR_THRESHOLD = 0.05 of every polybench verify(), e.g. polybench/gemm.cpp:96
    mfxCheck(mfx_percent_diff_count(d.handle(), rows, cols, X.get(), cols, R.get(), cols, 0.05,
                 &c.percent_diff_failures, d.stream()),
        "mfx_percent_diff_count");
    c.pass = c.percent_diff_failures == 0;
  } else {
    c.pass = std::isfinite(c.rel_frobenius) && c.rel_frobenius <= o.frobenius_tolerance(stages);
  }
  return c;
}

// Extended result columns shared by the contraction benchmarks.
inline void contractionExtras(std::map<std::string, std::string>& out, const ContractionOptions& o,
    const CheckResult& c) {
  std::ostringstream e;
  e.precision(3);
  e << std::scientific << c.rel_frobenius;
  out["precision"] = o.precision_name;
  out["rel-frobenius-error"] = e.str();
  out["percent-diff-failures"] = std::to_string(c.percent_diff_failures);
  out["init"] = o.init;
  out["seed"] = o.reference_inputs() ? "N/A" : std::to_string(o.seed);
}
