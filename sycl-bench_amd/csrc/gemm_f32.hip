// FP32 dense contractions on v_mfma_f32_32x32x2_f32.
//
// One workgroup computes a BM x BN tile of C with (BM/WM) x (BN/WN) waves;
// each wave owns a WM x WN sub-tile held as (WM/32) x (WN/32) 32x32
// accumulator blocks.  A and B panels of BK k-steps are staged in LDS with
// double buffering: the global loads of step kt+1 are issued before the MFMAs
// of step kt and land in the other LDS buffer afterwards, so one barrier per
// step suffices.  LDS holds both panels k-major (As[k][m], Bs[k][n]) whatever
// the global layout, padded by 4 floats per row so the transposing stores of
// k-contiguous operands hit distinct banks.
//
// Replaces the naive one-work-item-per-element kernels of
// polybench/gemm.cpp:77-92, polybench/2mm.cpp:94-123, polybench/3mm.cpp:115-158,
// polybench/syrk.cpp:70-84, polybench/syr2k.cpp:73-88 and
// runtime/matmulchain.cpp:15-29.
#include "mfx_internal.h"

namespace mfx {
namespace {

struct F32Params {
  DenseOp op;
  int tiles_m, tiles_n;
  int group;
  int vec_a, vec_b;  // 128-bit global loads allowed for A / B
};

template <int BM, int BN, int BK, int WM, int WN>
struct F32Cfg {
  static constexpr int NWM = BM / WM, NWN = BN / WN, NW = NWM * NWN, NT = NW * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int LDSA = BM + 4, LDSB = BN + 4;
};

// Loads the BK x ROWS panel of a k-major-in-LDS operand starting at (row0, k0)
// into registers.  kmajor: element (r, k) at P[r*ld + k]; else P[k*ld + r].
template <int ROWS, int BK, int NT, bool KMAJOR>
struct PanelLoader {
  static constexpr int NV = ROWS * BK / 4;  // float4 chunks
  static constexpr int PER = NV / NT;
  static_assert(NV % NT == 0, "panel must split evenly over threads");
  float4 v[PER];

  __device__ inline void load(const float* P, int64_t ld, int rows, int K, int row0, int k0, int tid, bool vec) {
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
          if(vec) {
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
          if(vec) {
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
  __device__ inline void store(float (*S)[LD], int tid) const {
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

template <int BM, int BN, int BK, int WM, int WN, bool AK, bool BKM>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_f32_kernel(const F32Params p) {
  using Cfg = F32Cfg<BM, BN, BK, WM, WN>;
  constexpr int NT = Cfg::NT, FM = Cfg::FM, FN = Cfg::FN;
  __shared__ float As[2][BK][Cfg::LDSA];
  __shared__ float Bs[2][BK][Cfg::LDSB];

  const DenseOp& op = p.op;
  int tm, tn;
  if(op.tiles == TILES_FULL) {
    const int pid = xcd_remap(blockIdx.x, gridDim.x);
    grouped_tile(pid, p.tiles_m, p.tiles_n, p.group, tm, tn);
  } else {
    const int pid = xcd_remap(blockIdx.x, gridDim.x);
    int ti, tj;
    lower_tile(pid, ti, tj);
    if(op.tiles == TILES_LOWER) { tm = ti; tn = tj; } else { tm = tj; tn = ti; }
  }
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
    const bool va = prod ? p.vec_b : p.vec_a, vb = prod ? p.vec_a : p.vec_b;
    la.load(pa, lda, op.M, op.K, m0, kt * BK, tid, va);
    lb.load(pb, ldb, op.N, op.K, n0, kt * BK, tid, vb);
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

struct F32Config {
  int bm, bn, bk, waves;
};
// Compiled tile configurations; index = config id.
constexpr F32Config kF32Configs[] = {
    {256, 256, 16, 4},  // 0: 128x128 per wave, large square problems
    {128, 128, 16, 4},  // 1: 64x64 per wave
    {64, 64, 32, 4},    // 2: 32x32 per wave, small problems
    {256, 128, 16, 4},  // 3: 128x64 per wave
};
constexpr int kF32ConfigCount = sizeof(kF32Configs) / sizeof(kF32Configs[0]);

template <int BM, int BN, int BK, int WM, int WN>
mfx_status launch_cfg(Handle* h, const F32Params& p, int grid, hipStream_t s) {
  using Cfg = F32Cfg<BM, BN, BK, WM, WN>;
  const bool ak = p.op.a_kmajor, bk = p.op.b_kmajor;
  return timed_launch(h, "gemm_f32_mfma", s, [&] {
    if(ak && !bk) gemm_f32_kernel<BM, BN, BK, WM, WN, true, false><<<grid, Cfg::NT, 0, s>>>(p);
    else if(ak && bk) gemm_f32_kernel<BM, BN, BK, WM, WN, true, true><<<grid, Cfg::NT, 0, s>>>(p);
    else if(!ak && !bk) gemm_f32_kernel<BM, BN, BK, WM, WN, false, false><<<grid, Cfg::NT, 0, s>>>(p);
    else gemm_f32_kernel<BM, BN, BK, WM, WN, false, true><<<grid, Cfg::NT, 0, s>>>(p);
  });
}

bool aligned16(const void* ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15u) == 0; }

}  // namespace

int f32_config_count() { return kF32ConfigCount; }
bool f32_config_info(int c, int* bm, int* bn, int* bk, int* waves) {
  if(c < 0 || c >= kF32ConfigCount) return false;
  *bm = kF32Configs[c].bm;
  *bn = kF32Configs[c].bn;
  *bk = kF32Configs[c].bk;
  *waves = kF32Configs[c].waves;
  return true;
}

mfx_status launch_dense_f32(Handle* h, const DenseOp& op, int config, hipStream_t s) {
  if(config < 0 || config >= kF32ConfigCount) return invalid("FP32 tile configuration out of range");
  F32Params p;
  p.op = op;
  const F32Config& c = kF32Configs[config];
  p.tiles_m = (int)ceil_div(op.M, c.bm);
  p.tiles_n = (int)ceil_div(op.N, c.bn);
  p.group = 8;
  // 128-bit loads need 16-byte aligned rows and a contiguous extent that is a
  // multiple of 4 (so a float4 is entirely inside or outside the matrix).
  auto vec_ok = [](const float* P, int64_t ld, bool kmajor, int rows, int K) {
    const int contig = kmajor ? K : rows;
    return aligned16(P) && ld % 4 == 0 && contig % 4 == 0;
  };
  p.vec_a = vec_ok(op.A, op.lda, op.a_kmajor, op.M, op.K);
  p.vec_b = vec_ok(op.B, op.ldb, op.b_kmajor, op.N, op.K);
  if(op.nprod == 2) {
    // second product swaps the operands: both must allow the same load width
    const bool v = p.vec_a && p.vec_b;
    p.vec_a = p.vec_b = v;
  }
  int grid;
  if(op.tiles == TILES_FULL) {
    grid = p.tiles_m * p.tiles_n;
  } else {
    if(p.tiles_m != p.tiles_n || c.bm != c.bn) return invalid("triangular schedule needs square tiles");
    grid = p.tiles_m * (p.tiles_m + 1) / 2;
  }
  if(grid == 0) return MFX_SUCCESS;
  switch(config) {
    case 0: return launch_cfg<256, 256, 16, 128, 128>(h, p, grid, s);
    case 1: return launch_cfg<128, 128, 16, 64, 64>(h, p, grid, s);
    case 2: return launch_cfg<64, 64, 32, 32, 32>(h, p, grid, s);
    case 3: return launch_cfg<256, 128, 16, 128, 64>(h, p, grid, s);
  }
  return invalid("unknown FP32 configuration");
}

}  // namespace mfx
