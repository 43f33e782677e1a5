// Roofline calibrators: the measured ceilings the contraction kernels are
// judged against.  The methodology follows micro/arith.cpp:41-68 (a
// register-only loop of dependent arithmetic whose result is known exactly)
// and micro/DRAM.cpp:47-62 (a device copy, read + write bytes per second),
// on the units that bound this path: the matrix cores and HBM.
//
// MFMA burn: every wave keeps NACC independent 32x32 accumulators and issues
// one MFMA into each per iteration, so the pipe never waits on a result; two
// workgroups of four waves per CU give two waves per SIMD.  All operands hold
// the same value, so after I iterations every accumulator element equals
// I * K * a * b exactly (powers of two), which the caller can check.
#include <algorithm>

#include "mfx_internal.h"

namespace mfx {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef int i32x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int NACC = 8;

// KIND 0: v_mfma_f32_32x32x2_f32 (K = 2); 1: v_mfma_f32_32x32x16_bf16 (K = 16);
// 2: v_mfma_scale_f32_32x32x64_f8f6f4 on e4m3 (K = 64).  `bits` carries the
// operand bit pattern so the compiler cannot fold the loop.
template <int KIND>
__global__ __launch_bounds__(256) void mfma_burn_kernel(uint32_t bits, int iters, float* out) {
  floatx16 acc[NACC];
#pragma unroll
  for(int i = 0; i < NACC; ++i)
#pragma unroll
    for(int q = 0; q < 16; ++q) acc[i][q] = 0.0f;
  if constexpr(KIND == 0) {
    const float a = __builtin_bit_cast(float, bits);
    for(int it = 0; it < iters; ++it)
#pragma unroll
      for(int i = 0; i < NACC; ++i) acc[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, a, acc[i], 0, 0, 0);
  } else if constexpr(KIND == 1) {
    const uint32_t w = (bits & 0xffffu) | (bits << 16);
    const bf16x8_t a = __builtin_bit_cast(bf16x8_t, (u32x4_t){w, w, w, w});
    for(int it = 0; it < iters; ++it)
#pragma unroll
      for(int i = 0; i < NACC; ++i) acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, a, acc[i], 0, 0, 0);
  } else {
    const int w = (int)(bits & 0xffu) * 0x01010101;
    const i32x8_t a = {w, w, w, w, w, w, w, w};
    for(int it = 0; it < iters; ++it)
#pragma unroll
      for(int i = 0; i < NACC; ++i)
        acc[i] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, a, acc[i], 0, 0, 0, 127, 0, 127);
  }
  float s = 0.0f;
#pragma unroll
  for(int i = 0; i < NACC; ++i) s += acc[i][0];
  if((threadIdx.x & 63) == 0) out[(blockIdx.x * blockDim.x + threadIdx.x) >> 6] = s;
}

// 128-bit copy of n chunks, one per thread over a grid that covers them all
// (measured faster on MI355X than grid-stride loops with several chunks in
// flight per thread: 6.25 against 4.4-5.0 TB/s for 2 GiB); the last `tail`
// (< 16) bytes after them are copied by the first thread.
__global__ __launch_bounds__(256) void hbm_copy_kernel(const f32x4_t* src, f32x4_t* dst, int64_t n, int tail) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if(i < n) dst[i] = src[i];
  if(tail && blockIdx.x == 0 && threadIdx.x == 0) {
    const unsigned char* s8 = reinterpret_cast<const unsigned char*>(src + n);
    unsigned char* d8 = reinterpret_cast<unsigned char*>(dst + n);
    for(int b = 0; b < tail; ++b) d8[b] = s8[b];
  }
}

}  // namespace
}  // namespace mfx

using namespace mfx;

extern "C" {

int64_t mfx_mfma_burn_waves(mfx_handle handle) {
  if(!handle) return 0;
  return (int64_t)std::max(1, handle->h.cu_count) * 2 * 4;
}

mfx_status mfx_mfma_burn(mfx_handle handle, mfx_precision prec, int64_t iterations, float* out, int64_t out_len,
    double* flop, float* expected, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  if(!out || !flop || !expected) return invalid("null output pointer");
  if(iterations <= 0 || iterations > (1 << 20)) return invalid("iterations must be in 1 .. 2^20");
  const int64_t waves = mfx_mfma_burn_waves(handle);
  if(out_len < waves) return invalid("out must hold mfx_mfma_burn_waves() floats");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int grid = (int)(waves / 4);
  const int it = (int)iterations;
  // Every product a*b*K is a power of two: 1/32, 1/4 or 64 per MFMA.
  switch(prec) {
    case MFX_PREC_FP32:
      *flop = (double)waves * NACC * iterations * 2.0 * 32 * 32 * 2;
      *expected = (float)(NACC * iterations) / 32.0f;
      return timed_launch(h, "mfma_burn_f32", s, [&] { mfma_burn_kernel<0><<<grid, 256, 0, s>>>(0x3e000000u, it, out); });
    case MFX_PREC_SPLIT_BF16:
    case MFX_PREC_BF16:
      *flop = (double)waves * NACC * iterations * 2.0 * 32 * 32 * 16;
      *expected = (float)(NACC * iterations) / 4.0f;
      return timed_launch(h, "mfma_burn_bf16", s, [&] { mfma_burn_kernel<1><<<grid, 256, 0, s>>>(0x3e00u, it, out); });
    case MFX_PREC_FP8_BLOCK:
    case MFX_PREC_FP8_BLOCK2:
      *flop = (double)waves * NACC * iterations * 2.0 * 32 * 32 * 64;
      *expected = (float)(NACC * iterations) * 64.0f;
      return timed_launch(h, "mfma_burn_fp8", s, [&] { mfma_burn_kernel<2><<<grid, 256, 0, s>>>(0x38u, it, out); });
  }
  return invalid("unknown precision");
}

mfx_status mfx_hbm_copy(mfx_handle handle, const void* src, void* dst, int64_t bytes, mfx_stream stream) {
  if(!handle) return invalid("null handle");
  Handle* h = &handle->h;
  MFX_HIP(hipSetDevice(h->device));
  if(bytes < 0) return invalid("negative byte count");
  if(bytes > 0 && (!src || !dst)) return invalid("null buffer");
  if((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u) return invalid("buffers must be 16-byte aligned");
  if(bytes == 0) return MFX_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = bytes / 16;
  const int tail = (int)(bytes % 16);
  if(n > (int64_t)0x7fffffff * 256) return invalid("copy larger than (2^31 - 1) * 256 chunks");
  const unsigned grid = (unsigned)std::max<int64_t>(1, ceil_div(n, 256));
  return timed_launch(h, "hbm_copy", s, [&] {
    hbm_copy_kernel<<<grid, 256, 0, s>>>(static_cast<const f32x4_t*>(src), static_cast<f32x4_t*>(dst), n, tail);
  });
}

}  // extern "C"
