// Entry points whose kernels are not built yet: they fail with
// MFX_ERROR_NOT_SUPPORTED and a message, never silently.
#include "mfx_internal.h"

namespace mfx {

int fp8_config_count() { return 0; }
bool fp8_config_info(int, int*, int*, int*, int*) { return false; }
size_t dense_fp8_workspace(const DenseOp&, bool, int) { return 0; }
mfx_status launch_dense_fp8(Handle*, const DenseOp&, bool, int, hipStream_t) {
  return unsupported("block-scaled FP8 contraction is not built yet");
}
mfx_status launch_quantize_fp8(Handle*, int64_t, int64_t, const float*, int64_t, int, uint8_t*, int32_t*, hipStream_t) {
  return unsupported("FP8 quantisation is not built yet");
}

}  // namespace mfx
