// Verification metric of the polybench benchmarks, as defined by the
// reference's polybench/common/polybenchUtilFuncts.h:14,27-42: two values
// pass when both are below 0.01 in magnitude, otherwise their difference is
// taken relative to the first (the expected value) in percent, with float
// narrowing through absVal.  The same rule runs on the GPU inside
// mfx_percent_diff_count for full matrices.
#pragma once

#define SMALL_FLOAT_VAL 0.00000001f

inline float absVal(float a) { return a < 0 ? (a * -1) : a; }

inline float percentDiff(double val1, double val2) {
  if((absVal(val1) < 0.01) && (absVal(val2) < 0.01)) return 0.0f;
  return 100.0f * (absVal(absVal(val1 - val2) / absVal(val1 + SMALL_FLOAT_VAL)));
}
