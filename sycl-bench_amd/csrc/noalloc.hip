// mfx_gemm_noalloc: mfx_gemm for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The packed operands of the split / single BF16 paths, which
// mfx_gemm keeps in the handle's growable workspace, live here in one
// file-scope device array of fixed size, `scratch`.
//
// Each device that loads this library holds its own copy of the array, at
// its own address, until the process ends.  That address is looked up once
// per device at set-up (noalloc_init, called by mfx_create with the device
// current) with hipGetSymbolAddress and kept in a per-device table; host
// code passes only that address to kernels.  (The array's name in host code
// is a host-side stand-in, not a device address.)
//
// Every call on one device shares that device's array: calls that overlap
// on one device (two streams, or two host threads) race.  The pack kernels
// overwrite every byte of both packed panels before the contraction reads
// them, so nothing has to be cleared between calls.
#include <mutex>

#include "mfx_internal.h"

namespace mfx {

mfx_status check_gemm_args(mfx_precision prec, mfx_op transA, mfx_op transB, int64_t m, int64_t n, int64_t k,
    const void* A, int64_t lda, const void* B, int64_t ldb, const void* C, int64_t ldc);
mfx_status dispatch_dense(Handle* h, mfx_precision prec, DenseOp& op, int config, hipStream_t s);

namespace {

static __device__ __align__(256) unsigned char scratch[MFX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

}  // namespace

// Looks up this device's copy of `scratch` (the device must be current).
mfx_status noalloc_init(int device) {
  if(device < 0 || device >= kMaxDevices) return unsupported("device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if(g_scratch[device]) return MFX_SUCCESS;
  void* p = nullptr;
  MFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if(!p) return hip_fail(hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
  g_scratch[device] = p;
  return MFX_SUCCESS;
}

static void* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_mutex);
  return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

}  // namespace mfx

using namespace mfx;

extern "C" {

mfx_status mfx_gemm_noalloc(mfx_handle handle, mfx_precision prec, mfx_op transA, mfx_op transB, int64_t m,
    int64_t n, int64_t k, float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C,
    int64_t ldc, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  MFX_CHECK(check_gemm_args(prec, transA, transB, m, n, k, A, lda, B, ldb, C, ldc));
  if(prec == MFX_PREC_FP8_BLOCK || prec == MFX_PREC_FP8_BLOCK2)
    return unsupported("mfx_gemm_noalloc: FP32, SPLIT_BF16 or BF16");
  if(m == 0 || n == 0) return MFX_SUCCESS;
  DenseOp op{};
  op.M = (int)m;
  op.N = (int)n;
  op.K = (int)k;
  op.A = A;
  op.lda = lda;
  op.a_kmajor = transA == MFX_OP_N;
  op.B = B;
  op.ldb = ldb;
  op.b_kmajor = transB == MFX_OP_T;
  op.nprod = 1;
  op.tiles = TILES_FULL;
  op.ep = Epilogue{C, ldc, alpha, beta, (int)m, (int)n, TILES_FULL, 0};
  const int cfg = choose_config(h, prec, m, n, k);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  // The kernels of this call are not recorded by mfx_timing_*: recording
  // could create events, which this entry point must not do.
  const bool timing = h->timing;
  h->timing = false;
  mfx_status st;
  if(prec == MFX_PREC_FP32) {
    // A and B are read directly; a split-K call keeps its partial tiles
    // (at most 32 MiB) in the fixed array
    void* ws = scratch_of(h->device);
    if(!ws) st = invalid("mfx_gemm_noalloc: no scratch address for this device (handle not made by mfx_create)");
    else st = launch_dense_f32_ws(h, op, cfg, ws, MFX_NOALLOC_SCRATCH_BYTES, s);
  } else {
    const bool split = prec == MFX_PREC_SPLIT_BF16;
    const size_t need = dense_bf16x_workspace(op, split, cfg);
    void* ws = scratch_of(h->device);
    if(need > MFX_NOALLOC_SCRATCH_BYTES) {
      st = unsupported("mfx_gemm_noalloc: the packed operands need " + std::to_string(need) +
                       " bytes; the fixed scratch array holds " + std::to_string((size_t)MFX_NOALLOC_SCRATCH_BYTES));
    } else if(!ws) {
      st = invalid("mfx_gemm_noalloc: no scratch address for this device (handle not made by mfx_create)");
    } else {
      st = launch_dense_bf16x_ws(h, op, split, cfg, ws, s);
    }
  }
  h->timing = timing;
  return st;
}

}  // extern "C"
