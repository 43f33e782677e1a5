// One BM x BN output tile of an FP32 contraction on v_mfma_f32_32x32x2_f32,
// shared by the single-problem kernel (gemm_f32.hip) and the persistent
// grouped kernel (grouped.hip).
//
// A workgroup of KG groups of (BM/WM) x (BN/WN) waves; each wave owns a
// WM x WN sub-tile held as (WM/32) x (WN/32) 32x32 accumulator blocks.  With
// KG > 1 every group computes the whole tile over its own 1/KG of each staged
// k block (more waves per CU for tiles that would otherwise leave the SIMDs
// waiting on loads), and the groups' accumulators are summed through LDS at
// the end, group 0 first, before group 0 writes the tile.  A and B panels of BK
// k-steps are staged in LDS with double buffering: the global loads of step
// s+1 are issued before the MFMAs of step s and land in the other LDS buffer
// afterwards, so one barrier per step suffices.  LDS holds both panels
// k-major (As[k][m], Bs[k][n]) whatever the global layout, padded by 4 floats
// per row so the transposing stores of k-contiguous operands hit distinct
// banks.
#pragma once

#include "mfx_internal.h"

namespace mfx {

template <int BM, int BN, int BK, int WM, int WN, int KG = 1>
struct F32Cfg {
  static constexpr int NWM = BM / WM, NWN = BN / WN, NWG = NWM * NWN, NW = NWG * KG, NT = NW * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int LDSA = BM + 4, LDSB = BN + 4;
  static constexpr int PANELS = 2 * BK * (LDSA + LDSB);
  static constexpr int REDUCE = (KG - 1) * BM * BN;  // accumulators of groups 1..KG-1
  static constexpr int SMEM_FLOATS = PANELS > REDUCE ? PANELS : REDUCE;
  static_assert(BK % (2 * KG) == 0, "each k-group takes whole MFMA k-steps of a stage");
};

// Loads the BK x ROWS panel of an operand starting at (row0, k0) into
// registers.  kmajor: element (r, k) at P[r*ld + k]; else P[k*ld + r].
// DEFER (128-bit loads only): load without a branch and zero the chunks
// outside the matrix at the LDS store instead (for two prefetch sets).
template <int ROWS, int BK, int NT, bool KMAJOR, bool DEFER = false>
struct PanelLoader {
  static constexpr int NV = ROWS * BK / 4;  // float4 chunks
  static constexpr int PER = NV / NT;
  static_assert(NV % NT == 0, "panel must split evenly over threads");
  static_assert(PER <= 32, "in-bounds flags are one bit per chunk");
  float4 v[PER];
  uint32_t in = 0;  // VEC and DEFER: bit i set when chunk i lies inside the matrix

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
      if constexpr(VEC && DEFER) {
        // Unconditional load from the nearest in-bounds chunk (rows >= 1 for
        // any scheduled tile, and the contiguous extent is a positive
        // multiple of 4 whenever a stage is issued); store() zeroes the
        // chunks outside the matrix.  With no branch around the load and its
        // value first used at the LDS store, the compiler can wait for one
        // prefetch set while a later one is still in flight.
        if(i == 0) in = 0;
        in |= (uint32_t)(gr < rows && gk < K) << i;
        const int cr = KMAJOR ? min(gr, rows - 1) : min(gr, rows - 4);
        const int ck = KMAJOR ? min(gk, K - 4) : min(gk, K - 1);
        v[i] = *reinterpret_cast<const float4*>(KMAJOR ? P + (int64_t)cr * ld + ck : P + (int64_t)ck * ld + cr);
      } else {
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
  }

  // S is the [BK][LD] k-major LDS panel; VEC as in load().
  template <int LD, bool VEC>
  __device__ __forceinline__ void store(float (*S)[LD], int tid) const {
#pragma unroll
    for(int i = 0; i < PER; ++i) {
      const int idx = tid + i * NT;
      float4 x = v[i];
      if constexpr(VEC && DEFER) {
        if(!((in >> i) & 1u)) x = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      if(KMAJOR) {
        const int r = idx / (BK / 4), k = (idx % (BK / 4)) * 4;
        S[k + 0][r] = x.x;
        S[k + 1][r] = x.y;
        S[k + 2][r] = x.z;
        S[k + 3][r] = x.w;
      } else {
        const int k = idx / (ROWS / 4), r = (idx % (ROWS / 4)) * 4;
        *reinterpret_cast<float4*>(&S[k][r]) = x;
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
template <int BM, int BN, int BK, int WM, int WN, bool AK, bool BKM, bool VEC, int KG = 1>
__device__ __forceinline__ void f32_tile(const DenseOp& op, int tm, int tn, float* smem) {
  using Cfg = F32Cfg<BM, BN, BK, WM, WN, KG>;
  constexpr int NT = Cfg::NT, FM = Cfg::FM, FN = Cfg::FN;
  // LDS row strides.  A k-contiguous operand is stored transposed, one float
  // at a time, by BK/4 threads per row: padding the row by 64/BK floats
  // (1 from BK = 64 up) puts the 4k-strided stores of those threads in
  // distinct banks.  Otherwise 128-bit stores need 16-byte rows (pad 4).
  constexpr int PT = BK >= 64 ? 1 : 64 / BK;
  constexpr int LA = AK ? BM + PT : Cfg::LDSA, LB = BKM ? BN + PT : Cfg::LDSB;
  static_assert(LA <= Cfg::LDSA && LB <= Cfg::LDSB, "panels fit the configuration's LDS");
  typedef float PanelA[BK][LA];
  typedef float PanelB[BK][LB];
  PanelA* As = reinterpret_cast<PanelA*>(smem);
  PanelB* Bs = reinterpret_cast<PanelB*>(smem + 2 * BK * LA);

  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int g = KG > 1 ? wid / Cfg::NWG : 0, w = wid - g * Cfg::NWG;  // k-group, wave within it
  const int wm = w / Cfg::NWN, wn = w % Cfg::NWN;

  floatx16 acc[FM][FN];
#pragma unroll
  for(int i = 0; i < FM; ++i)
#pragma unroll
    for(int j = 0; j < FN; ++j)
#pragma unroll
      for(int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  PanelLoader<BM, BK, NT, AK, (KG > 1)> la;
  PanelLoader<BN, BK, NT, BKM, (KG > 1)> lb;
  const int KT = (op.K + BK - 1) / BK;
  const int steps = KT * op.nprod;

  auto issue_to = [&](auto& xa, auto& xb, int s) __attribute__((always_inline)) {
    const int prod = s / KT, kt = s - prod * KT;
    const float* pa = prod ? op.B : op.A;
    const float* pb = prod ? op.A : op.B;
    const int64_t lda = prod ? op.ldb : op.lda, ldb = prod ? op.lda : op.ldb;
    xa.template load<VEC>(pa, lda, op.M, op.K, m0, kt * BK, tid);
    xb.template load<VEC>(pb, ldb, op.N, op.K, n0, kt * BK, tid);
  };
  auto commit_from = [&](const auto& xa, const auto& xb, int buf) __attribute__((always_inline)) {
    xa.template store<LA, VEC>(As[buf], tid);
    xb.template store<LB, VEC>(Bs[buf], tid);
  };
  auto issue = [&](int s) __attribute__((always_inline)) { issue_to(la, lb, s); };
  auto commit = [&](int buf) __attribute__((always_inline)) { commit_from(la, lb, buf); };
  const int lrow = lane & 31, lk = lane >> 5;
  constexpr int KS = BK / KG;  // k of each stage per group
  const int kg0 = g * KS;
  auto compute = [&](int cur) __attribute__((always_inline)) {
#pragma unroll
    for(int kk = 0; kk < KS; kk += 2) {
      float a[FM], b[FN];
#pragma unroll
      for(int i = 0; i < FM; ++i) a[i] = As[cur][kg0 + kk + lk][wm * WM + i * 32 + lrow];
#pragma unroll
      for(int j = 0; j < FN; ++j) b[j] = Bs[cur][kg0 + kk + lk][wn * WN + j * 32 + lrow];
#pragma unroll
      for(int i = 0; i < FM; ++i)
#pragma unroll
        for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };

  if constexpr(KG == 1) {
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
  } else {
    // One workgroup per CU: loads run two stages ahead, in two register
    // sets (la/lb for even stages, la1/lb1 for odd ones), so a stage's loads
    // have two stages of MFMAs to land.  At the top of the steady loop stage
    // s is in LDS buffer 0 and stage s+1 in flight in the odd set; the loop
    // body has no branches, so the compiler waits only for the older set.
    PanelLoader<BM, BK, NT, AK, true> la1;
    PanelLoader<BN, BK, NT, BKM, true> lb1;
    if(steps > 0) issue_to(la, lb, 0);
    if(steps > 1) issue_to(la1, lb1, 1);
    if(steps > 0) commit_from(la, lb, 0);
    __syncthreads();
    int s = 0;
    for(; s + 3 < steps; s += 2) {
      issue_to(la, lb, s + 2);
      compute(0);
      commit_from(la1, lb1, 1);
      __syncthreads();
      issue_to(la1, lb1, s + 3);
      compute(1);
      commit_from(la, lb, 0);
      __syncthreads();
    }
    const int left = steps - s;  // 0 (no k at all) to 3
    if(left == 3) {
      issue_to(la, lb, s + 2);
      compute(0);
      commit_from(la1, lb1, 1);
      __syncthreads();
      compute(1);
      commit_from(la, lb, 0);
      __syncthreads();
      compute(0);
    } else if(left == 2) {
      compute(0);
      commit_from(la1, lb1, 1);
      __syncthreads();
      compute(1);
    } else if(left == 1) {
      compute(0);
    }
  }

  if constexpr(KG > 1) {
    // Groups 1..KG-1 park their accumulators in LDS (over the panels, once
    // every group is done with them), lane-contiguous; group 0 adds them in
    // group order and alone writes the tile.
    __syncthreads();
    auto at = [&](int grp, int blk, int q) __attribute__((always_inline)) {
      return ((((grp - 1) * Cfg::NWG + w) * (FM * FN) + blk) * 16 + q) * 64 + lane;
    };
    if(g > 0) {
#pragma unroll
      for(int i = 0; i < FM; ++i)
#pragma unroll
        for(int j = 0; j < FN; ++j)
#pragma unroll
          for(int q = 0; q < 16; ++q) smem[at(g, i * FN + j, q)] = acc[i][j][q];
    }
    __syncthreads();
    if(g != 0) return;
    for(int gg = 1; gg < KG; ++gg)
#pragma unroll
      for(int i = 0; i < FM; ++i)
#pragma unroll
        for(int j = 0; j < FN; ++j)
#pragma unroll
          for(int q = 0; q < 16; ++q) acc[i][j][q] += smem[at(gg, i * FN + j, q)];
  }

  static_for<FM * FN>([&](auto ij) __attribute__((always_inline)) {
    constexpr int i = decltype(ij)::value / FN, j = decltype(ij)::value % FN;
    store_block32(op.ep, acc[i][j], m0 + wm * WM + i * 32, n0 + wn * WN + j * 32, lane);
  });
}

}  // namespace mfx
