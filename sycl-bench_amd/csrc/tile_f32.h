// One BM x BN output tile of an FP32 contraction on v_mfma_f32_32x32x2_f32,
// shared by the single-problem kernel (gemm_f32.hip) and the persistent
// grouped kernel (grouped.hip).
//
// A workgroup of (BM/WM) x (BN/WN) waves; each wave owns a WM x WN sub-tile
// held as (WM/32) x (WN/32) 32x32 accumulator blocks.  A and B panels of BK
// k-steps are staged in LDS with double buffering: the global loads of step
// s+1 are issued before the MFMAs of step s and land in the other LDS buffer
// afterwards, so one barrier per step suffices.  LDS holds both panels
// k-major (As[k][m], Bs[k][n]) whatever the global layout, padded by 4 floats
// per row so the transposing stores of k-contiguous operands hit distinct
// banks.
#pragma once

#include "mfx_internal.h"

namespace mfx {

template <int BM, int BN, int BK, int WM, int WN>
struct F32Cfg {
  static constexpr int NWM = BM / WM, NWN = BN / WN, NW = NWM * NWN, NT = NW * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int LDSA = BM + 4, LDSB = BN + 4;
  static constexpr int SMEM_FLOATS = 2 * BK * (LDSA + LDSB);
};

// Loads the BK x ROWS panel of an operand starting at (row0, k0) into
// registers.  kmajor: element (r, k) at P[r*ld + k]; else P[k*ld + r].
template <int ROWS, int BK, int NT, bool KMAJOR>
struct PanelLoader {
  static constexpr int NV = ROWS * BK / 4;  // float4 chunks
  static constexpr int PER = NV / NT;
  static_assert(NV % NT == 0, "panel must split evenly over threads");
  float4 v[PER];

  // VEC: 128-bit loads (compile-time, so the steady-state loop has no
  // per-load branches); otherwise four bounds-checked scalar loads.
  template <bool VEC>
  __device__ __forceinline__ void load(const float* P, int64_t ld, int rows, int K, int row0, int k0, int tid) {
#pragma unroll
    for(int i = 0; i < PER; ++i) {
      const int idx = tid + i * NT;
      int r, k;
      if(KMAJOR) {
        r = idx / (BK / 4);
        k = (idx % (BK / 4)) * 4;
      } else {
        k = idx / (ROWS / 4);
        r = (idx % (ROWS / 4)) * 4;
      }
      const int gr = row0 + r, gk = k0 + k;
      float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
      if(KMAJOR) {
        const float* src = P + (int64_t)gr * ld + gk;
        if(gr < rows) {
          if constexpr(VEC) {
            if(gk < K) x = *reinterpret_cast<const float4*>(src);
          } else {
            if(gk + 0 < K) x.x = src[0];
            if(gk + 1 < K) x.y = src[1];
            if(gk + 2 < K) x.z = src[2];
            if(gk + 3 < K) x.w = src[3];
          }
        }
      } else {
        const float* src = P + (int64_t)gk * ld + gr;
        if(gk < K) {
          if constexpr(VEC) {
            if(gr < rows) x = *reinterpret_cast<const float4*>(src);
          } else {
            if(gr + 0 < rows) x.x = src[0];
            if(gr + 1 < rows) x.y = src[1];
            if(gr + 2 < rows) x.z = src[2];
            if(gr + 3 < rows) x.w = src[3];
          }
        }
      }
      v[i] = x;
    }
  }

  // S is the [BK][LD] k-major LDS panel.
  template <int LD>
  __device__ __forceinline__ void store(float (*S)[LD], int tid) const {
#pragma unroll
    for(int i = 0; i < PER; ++i) {
      const int idx = tid + i * NT;
      if(KMAJOR) {
        const int r = idx / (BK / 4), k = (idx % (BK / 4)) * 4;
        S[k + 0][r] = v[i].x;
        S[k + 1][r] = v[i].y;
        S[k + 2][r] = v[i].z;
        S[k + 3][r] = v[i].w;
      } else {
        const int k = idx / (ROWS / 4), r = (idx % (ROWS / 4)) * 4;
        *reinterpret_cast<float4*>(&S[k][r]) = v[i];
      }
    }
  }
};

// 128-bit loads need 16-byte aligned rows and a contiguous extent that is a
// multiple of 4 (so a float4 is entirely inside or outside the matrix).
inline bool f32_vec_ok(const float* P, int64_t ld, bool kmajor, int rows, int K) {
  const int contig = kmajor ? K : rows;
  return (reinterpret_cast<uintptr_t>(P) & 15u) == 0 && ld % 4 == 0 && contig % 4 == 0;
}

// Computes tile (tm, tn) of `op` (all nprod products) and writes it through
// op.ep.  `smem` holds F32Cfg::SMEM_FLOATS floats; the caller separates two
// calls that reuse it with a barrier.
template <int BM, int BN, int BK, int WM, int WN, bool AK, bool BKM, bool VEC>
__device__ __forceinline__ void f32_tile(const DenseOp& op, int tm, int tn, float* smem) {
  using Cfg = F32Cfg<BM, BN, BK, WM, WN>;
  constexpr int NT = Cfg::NT, FM = Cfg::FM, FN = Cfg::FN;
  typedef float PanelA[BK][Cfg::LDSA];
  typedef float PanelB[BK][Cfg::LDSB];
  PanelA* As = reinterpret_cast<PanelA*>(smem);
  PanelB* Bs = reinterpret_cast<PanelB*>(smem + 2 * BK * Cfg::LDSA);

  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid / Cfg::NWN, wn = wid % Cfg::NWN;

  floatx16 acc[FM][FN];
#pragma unroll
  for(int i = 0; i < FM; ++i)
#pragma unroll
    for(int j = 0; j < FN; ++j)
#pragma unroll
      for(int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  PanelLoader<BM, BK, NT, AK> la;
  PanelLoader<BN, BK, NT, BKM> lb;
  const int KT = (op.K + BK - 1) / BK;
  const int steps = KT * op.nprod;

  auto issue = [&](int s) __attribute__((always_inline)) {
    const int prod = s / KT, kt = s - prod * KT;
    const float* pa = prod ? op.B : op.A;
    const float* pb = prod ? op.A : op.B;
    const int64_t lda = prod ? op.ldb : op.lda, ldb = prod ? op.lda : op.ldb;
    la.template load<VEC>(pa, lda, op.M, op.K, m0, kt * BK, tid);
    lb.template load<VEC>(pb, ldb, op.N, op.K, n0, kt * BK, tid);
  };
  auto commit = [&](int buf) __attribute__((always_inline)) {
    la.template store<Cfg::LDSA>(As[buf], tid);
    lb.template store<Cfg::LDSB>(Bs[buf], tid);
  };
  const int lrow = lane & 31, lk = lane >> 5;
  auto compute = [&](int cur) __attribute__((always_inline)) {
#pragma unroll
    for(int kk = 0; kk < BK; kk += 2) {
      float a[FM], b[FN];
#pragma unroll
      for(int i = 0; i < FM; ++i) a[i] = As[cur][kk + lk][wm * WM + i * 32 + lrow];
#pragma unroll
      for(int j = 0; j < FN; ++j) b[j] = Bs[cur][kk + lk][wn * WN + j * 32 + lrow];
#pragma unroll
      for(int i = 0; i < FM; ++i)
#pragma unroll
        for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };

  // Steady state without branches around the prefetch, so the registers
  // holding step s+1 stay live across the MFMAs of step s (no early wait).
  if(steps > 0) {
    issue(0);
    commit(0);
  }
  __syncthreads();
  for(int s = 0; s + 1 < steps; ++s) {
    issue(s + 1);
    compute(s & 1);
    commit((s + 1) & 1);
    __syncthreads();
  }
  if(steps > 0) compute((steps - 1) & 1);

  static_for<FM * FN>([&](auto ij) __attribute__((always_inline)) {
    constexpr int i = decltype(ij)::value / FN, j = decltype(ij)::value % FN;
    store_block32(op.ep, acc[i][j], m0 + wm * WM + i * 32, n0 + wn * WN + j * 32, lane);
  });
}

}  // namespace mfx
