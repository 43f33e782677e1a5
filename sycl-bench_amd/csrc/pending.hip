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

using namespace mfx;

extern "C" {

mfx_status mfx_gemm_grouped(mfx_handle, mfx_precision, const mfx_gemm_problem*, int32_t, const mfx_dep*, int32_t,
    mfx_stream) {
  return unsupported("grouped launch is not built yet");
}
mfx_status mfx_chain(mfx_handle, mfx_precision, int32_t, const float* const*, int64_t, float* const*, float*,
    mfx_stream) {
  return unsupported("chain launch is not built yet");
}
mfx_status mfx_grouped_graph_create(mfx_handle, mfx_precision, const mfx_gemm_problem*, int32_t, const mfx_dep*,
    int32_t, mfx_stream, mfx_graph*) {
  return unsupported("grouped graph is not built yet");
}
mfx_status mfx_graph_launch(mfx_graph, mfx_stream) { return unsupported("grouped graph is not built yet"); }
mfx_status mfx_graph_destroy(mfx_graph) { return MFX_SUCCESS; }
mfx_status mfx_autotune(mfx_handle, mfx_precision, int64_t, int64_t, int64_t, int32_t*, mfx_stream) {
  return unsupported("autotuner is not built yet");
}
mfx_status mfx_autotune_save(const char*) { return unsupported("autotuner is not built yet"); }
mfx_status mfx_autotune_load(const char*) { return unsupported("autotuner is not built yet"); }

}  // extern "C"
