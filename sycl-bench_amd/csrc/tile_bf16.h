// Split-BF16 / single-BF16 building blocks shared by the single-problem
// kernels (gemm_bf16.hip) and the persistent grouped kernel (grouped.hip).
//
// Split BF16: every FP32 operand x is written as hi + lo with
// hi = bf16_rne(x), lo = bf16_rne(x - hi); the product uses
// lo*hi + hi*lo + hi*hi (the lo*lo term, ~2^-16 relative, is dropped) with
// FP32 accumulation.  Single BF16 keeps only hi*hi.
//
// Packed "fragment order" layout of an operand (rows x K, rows = M for the A
// role, N for the B role):
//     P[blk][kb][hl][lane][8]   (bf16; one 16-byte chunk per lane)
// blk = 32-row block, kb = 16-wide k block, hl = hi / lo, and lane l holds
// row blk*32 + l%32, k = kb*16 + 8*(l/32) + 0..7 -- the register operand of
// v_mfma_f32_32x32x16_bf16 for both the A and the B role.  Rows are padded to
// the tile size and k to the 32-wide stage with zeros.
#pragma once

#include "mfx_internal.h"

namespace mfx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// Native clang vector (not HIP's uint4 struct) so register arrays of chunks are
// promoted to registers instead of living in scratch memory.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// 16-wide k blocks per LDS stage: 2 (BK = 32) for split BF16, whose hi/lo
// planes fill a 64 KiB stage of a 256x256 tile; 4 (BK = 64) for single BF16,
// so its stages carry as many bytes and MFMAs per barrier.
template <int NHL>
constexpr int kb_per_stage() {
  return NHL == 2 ? 2 : 4;
}
__host__ __device__ inline int kb_per_stage_of(int nhl) { return nhl == 2 ? 2 : 4; }

// bf16_tile keeps each thread's chunk offsets within a tile's panel in 32
// bits: at most (BM / 32) blocks of kbs * NHL * 64 chunks, BM <= 256.  So the
// packed paths take K only while 8 * kbs * NHL * 64 < 2^31, i.e. K below
// 2^25 (split BF16) or 2^26 (single BF16).
inline bool bf16_k_supported(int64_t k, int nhl) {
  const int64_t stage = 16 * kb_per_stage_of(nhl);
  const int64_t kbs = (k + stage - 1) / stage * stage / 16;
  return 8 * kbs * nhl * 64 < ((int64_t)1 << 31);
}

__device__ __forceinline__ uint32_t pack_bf16_pair(uint16_t a, uint16_t b) {
  return (uint32_t)a | ((uint32_t)b << 16);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two floats to a packed pair of BF16, round to nearest even, in one
// instruction (v_cvt_pk_bf16_f32); equals bf16_rne on every non-NaN input.
__device__ __forceinline__ uint32_t cvt_pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// Splits 8 floats into their BF16 hi parts (and, for NHL == 2, the BF16 lo
// parts of the residuals), packed as one 16-byte chunk each.
template <int NHL>
__device__ __forceinline__ void split_chunk(const float (&v)[8], u32x4& h, u32x4& l) {
  uint32_t hw[4], lw[4];
#pragma unroll
  for(int p = 0; p < 4; ++p) {
    hw[p] = cvt_pk_bf16(v[2 * p], v[2 * p + 1]);
    if(NHL == 2) {
      const float h0 = __builtin_bit_cast(float, hw[p] << 16);
      const float h1 = __builtin_bit_cast(float, hw[p] & 0xffff0000u);
      lw[p] = cvt_pk_bf16(v[2 * p] - h0, v[2 * p + 1] - h1);
    } else {
      lw[p] = 0;
    }
  }
  h = (u32x4){hw[0], hw[1], hw[2], hw[3]};
  l = (u32x4){lw[0], lw[1], lw[2], lw[3]};
}

// Source view of one operand for packing.
struct PackSrc {
  const float* src;
  int64_t ld;
  int rows, K;  // logical extent (rows x K)
  int kmajor;   // element (r, k) at src[r*ld + k], else src[k*ld + r]
  int vec;      // 128-bit loads allowed (kmajor, aligned rows)
};

// Host-side constructor (decides whether 128-bit loads are allowed).
PackSrc make_pack_src(const float* src, int64_t ld, int rows, int K, bool kmajor);

// Loads lane-chunk (blk, kb, lane) of an operand: 8 consecutive k of row
// blk*32 + lane%32 starting at k = kb*16 + 8*(lane/32); zeros outside it.
__device__ __forceinline__ void load_chunk8(const PackSrc& s, int blk, int kb, int lane, float (&v)[8]) {
  const int r = blk * 32 + (lane & 31);
  const int k0 = kb * 16 + (lane >> 5) * 8;
  if(s.kmajor) {
    const float* p = s.src + (int64_t)r * s.ld + k0;
    if(r < s.rows && s.vec && k0 + 8 <= s.K) {
      const float4 x = reinterpret_cast<const float4*>(p)[0];
      const float4 y = reinterpret_cast<const float4*>(p)[1];
      v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
      v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    } else {
#pragma unroll
      for(int e = 0; e < 8; ++e) v[e] = (r < s.rows && k0 + e < s.K) ? p[e] : 0.0f;
    }
  } else {
#pragma unroll
    for(int e = 0; e < 8; ++e) v[e] = (r < s.rows && k0 + e < s.K) ? s.src[(int64_t)(k0 + e) * s.ld + r] : 0.0f;
  }
}

// LDS floats per wave for load_chunk8_staged: 16 k rows of 32 floats, rows
// 8..15 shifted by 32 floats so the two half-waves' column reads fall in
// different banks.
constexpr int STAGED_TILE_FLOATS = 16 * 32 + 32;

// load_chunk8 for a non-k-major source (element (r, k) at src[k*ld + r]),
// staged through LDS: the 64 lanes of one wave hold the 64 lane-chunks of one
// (blk, kb) pair; together they load its 16 x 32 (k x row) source tile as
// 16 row segments of 128 bytes (128-bit loads when s.vec), then each lane
// reads its 8 k values of one row from LDS.  `tile` is the wave's
// STAGED_TILE_FLOATS area.  Contains a workgroup barrier: every thread of the
// block calls it, `active` false for threads without a chunk (whose wave may
// still hold active lanes).
__device__ __forceinline__ void load_chunk8_staged(const PackSrc& s, int blk, int kb, int lane, bool active,
    float (&v)[8], float* tile) {
  const int r0 = blk * 32, k0 = kb * 16;
  if(active) {
#pragma unroll
    for(int h = 0; h < 2; ++h) {
      const int q = lane + h * 64;  // 128 float4 slots: 16 k rows x 8
      const int k = q >> 3, r = (q & 7) * 4;
      const int gk = k0 + k, gr = r0 + r;
      float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if(gk < s.K) {
        const float* p = s.src + (int64_t)gk * s.ld + gr;
        if(s.vec && gr + 4 <= s.rows) {
          x = *reinterpret_cast<const float4*>(p);
        } else {
          if(gr < s.rows) x.x = p[0];
          if(gr + 1 < s.rows) x.y = p[1];
          if(gr + 2 < s.rows) x.z = p[2];
          if(gr + 3 < s.rows) x.w = p[3];
        }
      }
      *reinterpret_cast<float4*>(tile + k * 32 + (k >= 8 ? 32 : 0) + r) = x;
    }
  }
  __syncthreads();
  if(active) {
    const int kh = (lane >> 5) * 8, r = lane & 31;
#pragma unroll
    for(int e = 0; e < 8; ++e) v[e] = tile[(kh + e) * 32 + (kh ? 32 : 0) + r];
  }
}

// load_chunk8_staged in two halves, so that a caller can keep the global
// loads of several chunks in flight: staged_load issues the lane's two
// 128-bit loads of the (blk, kb) pair's 16 x 32 source tile into registers;
// staged_read, after the caller has written them to the wave's LDS tile with
// staged_write and synchronised, reads the lane's 8 k values of its row.
__device__ __forceinline__ void staged_load(const PackSrc& s, int blk, int kb, int lane, bool active, float4 (&x)[2]) {
  const int r0 = blk * 32, k0 = kb * 16;
#pragma unroll
  for(int h = 0; h < 2; ++h) {
    const int q = lane + h * 64;
    const int gk = k0 + (q >> 3), gr = r0 + (q & 7) * 4;
    x[h] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if(active && gk < s.K) {
      const float* p = s.src + (int64_t)gk * s.ld + gr;
      if(s.vec && gr + 4 <= s.rows) {
        x[h] = *reinterpret_cast<const float4*>(p);
      } else {
        if(gr < s.rows) x[h].x = p[0];
        if(gr + 1 < s.rows) x[h].y = p[1];
        if(gr + 2 < s.rows) x[h].z = p[2];
        if(gr + 3 < s.rows) x[h].w = p[3];
      }
    }
  }
}
__device__ __forceinline__ void staged_write(const float4 (&x)[2], int lane, float* tile) {
#pragma unroll
  for(int h = 0; h < 2; ++h) {
    const int q = lane + h * 64, k = q >> 3;
    *reinterpret_cast<float4*>(tile + k * 32 + (k >= 8 ? 32 : 0) + (q & 7) * 4) = x[h];
  }
}
__device__ __forceinline__ void staged_read(int lane, const float* tile, float (&v)[8]) {
  const int kh = (lane >> 5) * 8, r = lane & 31;
#pragma unroll
  for(int e = 0; e < 8; ++e) v[e] = tile[(kh + e) * 32 + (kh ? 32 : 0) + r];
}

// Splits a loaded lane-chunk and stores it at its fragment-order position in
// the packed operand `dst` (kbs = number of 16-wide k blocks).
template <int NHL>
__device__ __forceinline__ void store_chunk(u32x4* dst, int blk, int kb, int lane, int kbs, const float (&v)[8]) {
  const int64_t base = ((int64_t)blk * kbs + kb) * NHL * 64 + lane;
  u32x4 h, l;
  split_chunk<NHL>(v, h, l);
  dst[base] = h;
  if(NHL == 2) dst[base + 64] = l;
}

// Packs lane-chunk (blk, kb, lane): 8 consecutive k of one row, hi (and lo).
template <int NHL>
__device__ __forceinline__ void pack_chunk(const PackSrc& s, int blk, int kb, int lane, int kbs, u32x4* dst) {
  float v[8];
  load_chunk8(s, blk, kb, lane, v);
  store_chunk<NHL>(dst, blk, kb, lane, kbs, v);
}

template <int BM, int BN, int WM, int WN, int NHL>
struct BfCfg {
  static constexpr int NWM = BM / WM, NWN = BN / WN, NW = NWM * NWN, NT = NW * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int KB = kb_per_stage<NHL>();
  static constexpr int CH_A = BM / 32 * KB * NHL * 64;  // 16-byte chunks per stage
  static constexpr int CH_B = BN / 32 * KB * NHL * 64;
  static constexpr int PER_A = CH_A / NT, PER_B = CH_B / NT;
  static constexpr int SMEM_CHUNKS = 2 * (CH_A + CH_B);
  static_assert(CH_A % NT == 0 && CH_B % NT == 0, "stage must split evenly");
};

// Packed operands of one contraction.
struct BfTileArgs {
  const u32x4* PA;  // packed A-role operand of product 0
  const u32x4* PB;  // packed B-role operand of product 0
  int kbs;          // k blocks of 16 in both packed operands
  int nblk_a;       // 32-row blocks held by PA / PB (rows padded to 32 only)
  int nblk_b;
  int nprod;        // 2: second product uses (PB, PA)
  Epilogue ep;
};

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 u) { return __builtin_bit_cast(bf16x8, u); }

// MFMAs of one 16-wide k block `kb` of an LDS stage: acc[i][j] +=
// A-fragments x B-fragments; sa / sb are the stage's A / B chunks.
template <int FM, int FN, int NHL>
__device__ __forceinline__ void bf16_kb_mma(floatx16 (&acc)[FM][FN], const u32x4* sa, const u32x4* sb, int wm, int wn,
    int lane, int kb) {
  bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
  for(int i = 0; i < FM; ++i) {
    const int blk = wm * FM + i;
    ah[i] = as_bf16x8(sa[(blk * kb_per_stage<NHL>() + kb) * NHL * 64 + lane]);
    if(NHL == 2) al[i] = as_bf16x8(sa[((blk * kb_per_stage<NHL>() + kb) * NHL + 1) * 64 + lane]);
  }
#pragma unroll
  for(int j = 0; j < FN; ++j) {
    const int blk = wn * FN + j;
    bh[j] = as_bf16x8(sb[(blk * kb_per_stage<NHL>() + kb) * NHL * 64 + lane]);
    if(NHL == 2) bl[j] = as_bf16x8(sb[((blk * kb_per_stage<NHL>() + kb) * NHL + 1) * 64 + lane]);
  }
  // Product-major order: consecutive MFMAs write different accumulators, so
  // none waits on the previous one's result.  The small cross terms go in
  // first so they are added while the accumulator is smallest.
  if(NHL == 2) {
#pragma unroll
    for(int i = 0; i < FM; ++i)
#pragma unroll
      for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for(int i = 0; i < FM; ++i)
#pragma unroll
      for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for(int i = 0; i < FM; ++i)
#pragma unroll
    for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
}

// bf16_kb_mma for a stage of KBS k blocks (chunk (blk, kb, hl, lane) of a
// stage at ((blk * KBS + kb) * NHL + hl) * 64 + lane).
template <int FM, int FN, int NHL, int KBS>
__device__ __forceinline__ void bf16_kb_mma_n(floatx16 (&acc)[FM][FN], const u32x4* sa, const u32x4* sb, int wm, int wn,
    int lane, int kb) {
  bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
  for(int i = 0; i < FM; ++i) {
    const int blk = wm * FM + i;
    ah[i] = as_bf16x8(sa[(blk * KBS + kb) * NHL * 64 + lane]);
    if(NHL == 2) al[i] = as_bf16x8(sa[((blk * KBS + kb) * NHL + 1) * 64 + lane]);
  }
#pragma unroll
  for(int j = 0; j < FN; ++j) {
    const int blk = wn * FN + j;
    bh[j] = as_bf16x8(sb[(blk * KBS + kb) * NHL * 64 + lane]);
    if(NHL == 2) bl[j] = as_bf16x8(sb[((blk * KBS + kb) * NHL + 1) * 64 + lane]);
  }
  if(NHL == 2) {
#pragma unroll
    for(int i = 0; i < FM; ++i)
#pragma unroll
      for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for(int i = 0; i < FM; ++i)
#pragma unroll
      for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for(int i = 0; i < FM; ++i)
#pragma unroll
    for(int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
}

// MFMAs of a whole LDS stage (kb_per_stage k blocks).
template <int FM, int FN, int NHL>
__device__ __forceinline__ void bf16_stage_mma(floatx16 (&acc)[FM][FN], const u32x4* sa, const u32x4* sb, int wm,
    int wn, int lane) {
#pragma unroll
  for(int kb = 0; kb < kb_per_stage<NHL>(); ++kb) bf16_kb_mma<FM, FN, NHL>(acc, sa, sb, wm, wn, lane, kb);
}

template <int FM, int FN>
__device__ __forceinline__ void zero_acc(floatx16 (&acc)[FM][FN]) {
#pragma unroll
  for(int i = 0; i < FM; ++i)
#pragma unroll
    for(int j = 0; j < FN; ++j)
#pragma unroll
      for(int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;
}

// Computes tile (tm, tn) from packed operands and writes it through t.ep.
// `lds` holds BfCfg::SMEM_CHUNKS chunks; the caller separates two calls that
// reuse it with a barrier.
template <int BM, int BN, int WM, int WN, int NHL>
__device__ __forceinline__ void bf16_tile(const BfTileArgs& t, int tm, int tn, u32x4* lds) {
  using Cfg = BfCfg<BM, BN, WM, WN, NHL>;
  constexpr int NT = Cfg::NT, FM = Cfg::FM, FN = Cfg::FN;
  constexpr int SA = Cfg::CH_A, SB = Cfg::CH_B;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid / Cfg::NWN, wn = wid % Cfg::NWN;

  floatx16 acc[FM][FN];
  zero_acc(acc);

  // Chunk c of a stage: block c / RUN of the tile, then the contiguous
  // (kb, hl, lane) run of that block for this stage.
  constexpr int RUN = Cfg::KB * NHL * 64;
  const int KT = t.kbs / Cfg::KB;
  const int steps = KT * t.nprod;
  const int64_t blk_stride = (int64_t)t.kbs * NHL * 64;
  // Packed operands hold only the 32-row blocks that meet the matrix: a tile
  // block past the last one reads the last block again.  Those rows feed only
  // output rows / columns past M / N, which the epilogue never stores, so any
  // finite data will do.
  // Global -> LDS without registers (global_load_lds_dwordx4): for one wave
  // and one i, the 64 lanes' chunks c = wave*64 + i*NT + lane are 64
  // consecutive chunks both in the packed operand (RUN is a multiple of 64)
  // and in the LDS stage, so each instruction moves one contiguous 1 KiB.
  // The LDS base of a wave is uniform: readfirstlane lets the compiler keep it
  // (and so M0) in scalar registers.  Per-thread chunk offsets within the
  // tile's A and B panels are computed once (32-bit: at most BM/32 blocks of
  // one panel); a stage only moves the wave-uniform panel base along k.
  const int wave_base = __builtin_amdgcn_readfirstlane(tid - lane);
  int off_a[2][Cfg::PER_A], off_b[2][Cfg::PER_B];
#pragma unroll
  for(int p = 0; p < 2; ++p) {
    const int last_a = (p ? t.nblk_b : t.nblk_a) - 1, last_b = (p ? t.nblk_a : t.nblk_b) - 1;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i) {
      const int c = tid + i * NT;
      const int blk = c / RUN, run = c - blk * RUN;
      off_a[p][i] = (min(tm * (BM / 32) + blk, last_a) - tm * (BM / 32)) * (int)blk_stride + run;
    }
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i) {
      const int c = tid + i * NT;
      const int blk = c / RUN, run = c - blk * RUN;
      off_b[p][i] = (min(tn * (BN / 32) + blk, last_b) - tn * (BN / 32)) * (int)blk_stride + run;
    }
  }
  auto issue = [&](int s, int buf) __attribute__((always_inline)) {
    const int prod = s / KT, kt = s - prod * KT;
    const u32x4* A = (prod ? t.PB : t.PA) + ((int64_t)tm * (BM / 32) * blk_stride + (int64_t)kt * RUN);
    const u32x4* B = (prod ? t.PA : t.PB) + ((int64_t)tn * (BN / 32) * blk_stride + (int64_t)kt * RUN);
    u32x4* sa = lds + buf * (SA + SB);
    u32x4* sb = sa + SA;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i) lds_dma16(A + (prod ? off_a[1][i] : off_a[0][i]), sa + wave_base + i * NT);
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i) lds_dma16(B + (prod ? off_b[1][i] : off_b[0][i]), sb + wave_base + i * NT);
  };
  auto compute = [&](int cur) __attribute__((always_inline)) {
    const u32x4* sa = lds + cur * (SA + SB);
    bf16_stage_mma<FM, FN, NHL>(acc, sa, sa + SA, wm, wn, lane);
  };

  // Stage s+1 streams into the other buffer while stage s is contracted; the
  // buffer it overwrites was last read before the previous barrier.
  if(steps > 0) issue(0, 0);
  lds_dma_wait();
  __syncthreads();
  for(int s = 0; s + 1 < steps; ++s) {
    issue(s + 1, (s + 1) & 1);
    compute(s & 1);
    lds_dma_wait();
    __syncthreads();
  }
  if(steps > 0) compute((steps - 1) & 1);

  static_for<FM * FN>([&](auto ij) __attribute__((always_inline)) {
    constexpr int i = decltype(ij)::value / FN, j = decltype(ij)::value % FN;
    store_block32(t.ep, acc[i][j], tm * BM + wm * WM + i * 32, tn * BN + wn * WN + j * 32, lane);
  });
}

// Waits until at most n of this wave's vector memory operations are
// outstanding (s_waitcnt vmcnt(n), n <= 63; expcnt and lgkmcnt left alone).
template <int n>
__device__ __forceinline__ void vm_wait() {
  static_assert(n >= 0 && n < 64, "vmcnt holds 6 bits");
  __builtin_amdgcn_s_waitcnt((n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8));
}

// bf16_tile with a deeper pipeline: NB LDS buffers of KB k blocks each, NB-1
// stages in flight.  A tile whose panels come from HBM rather than L2 (one
// tile per problem, as in a batch of small problems) is bound by the bytes
// it keeps in flight: two 64 KiB stages (one in flight) stream at about
// 3.7 TB/s over the chip, so four 32 KiB stages (three in flight) in the
// same LDS double the bytes in flight.  The MFMA sequence -- k blocks in
// order, products in the same order -- is that of bf16_tile, so both give
// bit-identical results.  `lds` holds NB * (CH_A + CH_B) chunks.
template <int BM, int BN, int WM, int WN, int NHL, int KB, int NB>
struct BfPipeCfg {
  static constexpr int NWM = BM / WM, NWN = BN / WN, NW = NWM * NWN, NT = NW * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int CH_A = BM / 32 * KB * NHL * 64;  // 16-byte chunks per stage
  static constexpr int CH_B = BN / 32 * KB * NHL * 64;
  static constexpr int PER_A = CH_A / NT, PER_B = CH_B / NT;
  static constexpr int SMEM_CHUNKS = NB * (CH_A + CH_B);
  static_assert(CH_A % NT == 0 && CH_B % NT == 0, "stage must split evenly");
  static_assert(kb_per_stage<NHL>() % KB == 0, "stages divide the packed k padding");
  static_assert(NB >= 2 && (NB - 2) * (PER_A + PER_B) < 64, "in-flight DMA count fits vmcnt");
};

template <int BM, int BN, int WM, int WN, int NHL, int KB, int NB>
__device__ __forceinline__ void bf16_tile_pipe(const BfTileArgs& t, int tm, int tn, u32x4* lds) {
  using Cfg = BfPipeCfg<BM, BN, WM, WN, NHL, KB, NB>;
  constexpr int NT = Cfg::NT, FM = Cfg::FM, FN = Cfg::FN;
  constexpr int SA = Cfg::CH_A, SB = Cfg::CH_B;
  constexpr int PER = Cfg::PER_A + Cfg::PER_B;  // DMA instructions per stage and wave
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid / Cfg::NWN, wn = wid % Cfg::NWN;

  floatx16 acc[FM][FN];
  zero_acc(acc);

  constexpr int RUN = KB * NHL * 64;
  const int KT = t.kbs / KB;
  const int steps = KT * t.nprod;
  const int64_t blk_stride = (int64_t)t.kbs * NHL * 64;
  const int wave_base = __builtin_amdgcn_readfirstlane(tid - lane);
  int off_a[2][Cfg::PER_A], off_b[2][Cfg::PER_B];
#pragma unroll
  for(int p = 0; p < 2; ++p) {
    const int last_a = (p ? t.nblk_b : t.nblk_a) - 1, last_b = (p ? t.nblk_a : t.nblk_b) - 1;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i) {
      const int c = tid + i * NT;
      const int blk = c / RUN, run = c - blk * RUN;
      off_a[p][i] = (min(tm * (BM / 32) + blk, last_a) - tm * (BM / 32)) * (int)blk_stride + run;
    }
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i) {
      const int c = tid + i * NT;
      const int blk = c / RUN, run = c - blk * RUN;
      off_b[p][i] = (min(tn * (BN / 32) + blk, last_b) - tn * (BN / 32)) * (int)blk_stride + run;
    }
  }
  auto issue = [&](int s) __attribute__((always_inline)) {
    const int prod = s / KT, kt = s - prod * KT;
    const u32x4* A = (prod ? t.PB : t.PA) + ((int64_t)tm * (BM / 32) * blk_stride + (int64_t)kt * RUN);
    const u32x4* B = (prod ? t.PA : t.PB) + ((int64_t)tn * (BN / 32) * blk_stride + (int64_t)kt * RUN);
    u32x4* sa = lds + (s % NB) * (SA + SB);
    u32x4* sb = sa + SA;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i) lds_dma16(A + (prod ? off_a[1][i] : off_a[0][i]), sa + wave_base + i * NT);
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i) lds_dma16(B + (prod ? off_b[1][i] : off_b[0][i]), sb + wave_base + i * NT);
  };

  // Stages 0 .. NB-2 in flight before the loop; iteration s waits for
  // stage s (its own DMA, then the barrier for everyone's), which also means
  // every wave has finished contracting stage s-1, so stage s+NB-1 can reuse
  // that buffer.
#pragma unroll
  for(int p = 0; p < NB - 1; ++p)
    if(p < steps) issue(p);
  for(int s = 0; s < steps; ++s) {
    const int after = steps - 1 - s;  // stages issued after s, at most NB - 2 in flight
    if(after >= NB - 2) vm_wait<(NB - 2) * PER>();
    else if(NB > 3 && after == 1) vm_wait<PER>();
    else lds_dma_wait();
    __syncthreads();
    if(s + NB - 1 < steps) issue(s + NB - 1);
    const u32x4* sa = lds + (s % NB) * (SA + SB);
#pragma unroll
    for(int kb = 0; kb < KB; ++kb) bf16_kb_mma_n<FM, FN, NHL, KB>(acc, sa, sa + SA, wm, wn, lane, kb);
  }

  static_for<FM * FN>([&](auto ij) __attribute__((always_inline)) {
    constexpr int i = decltype(ij)::value / FN, j = decltype(ij)::value % FN;
    store_block32(t.ep, acc[i][j], tm * BM + wm * WM + i * 32, tn * BN + wn * WN + j * 32, lane);
  });
}

// ---------------------------------------------------------------------------
// Direct tile: the FP32 operands go from global memory straight into LDS
// (global_load_lds_dwordx4, no registers) and each wave splits its own
// fragments into hi / lo BF16 right before its MFMAs -- no packing pass and
// no packed copy in HBM.  The split is split_chunk's, applied to the same 8
// consecutive k of the same row, and the MFMA sequence is bf16_tile_pipe's,
// so the result is bit-identical to the packed path.
//
// One LDS stage holds 16 k of both panels; a 32-row (or 32-column) block of
// an operand is 128 16-byte chunks laid out so that one DMA instruction of a
// wave fills 64 consecutive chunks:
//   k-major source (element (r, k) at p[r*ld + k]): chunk h*64 + l of a
//     block holds row l%32, k = 8*(l/32) + 4h .. +3, so a lane's fragment
//     (8 consecutive k of its row) is chunks l and 64 + l;
//   row-major source (element (r, k) at p[k*ld + r]): the block is a 16 x 32
//     (k x r) float array, row k at floats k*32 .. k*32+31, filled 4
//     consecutive r per chunk, and a lane reads its 8 k one float each.
// Rows past the operand read its last row (or last 4 columns) again; they
// only feed outputs the epilogue does not store.  Requirements (checked by
// the host): K a multiple of 16; 16-byte aligned operands with ld % 4 == 0;
// for a row-major source also rows % 4 == 0.
struct DirectOp {
  const float* p;
  int64_t ld;
  int rows;    // logical rows of the operand's role (M for A, N for B)
  int kmajor;
};
struct DirectTileArgs {
  DirectOp a, b;
  int K;
  Epilogue ep;
};

template <int BM, int BN, int WM, int WN, int NB>
struct BfDirectCfg {
  static constexpr int NWM = BM / WM, NWN = BN / WN, NW = NWM * NWN, NT = NW * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int CH_A = BM / 32 * 128, CH_B = BN / 32 * 128;  // 16-byte chunks per stage
  static constexpr int PER_A = CH_A / NT, PER_B = CH_B / NT;
  static constexpr int SMEM_CHUNKS = NB * (CH_A + CH_B);
  static_assert(CH_A % NT == 0 && CH_B % NT == 0, "stage must split evenly");
  static_assert(NB >= 2 && (NB - 2) * (PER_A + PER_B) < 64, "in-flight DMA count fits vmcnt");
};

// Source of DMA group g (64 chunks: block g / 2, half g % 2) for lane l of a
// tile whose operand rows start at row0, at k = 0.
__device__ __forceinline__ const float* direct_src(const DirectOp& o, int row0, int g, int l) {
  const int blk = g >> 1, h = g & 1;
  if(o.kmajor) {
    const int r = min(row0 + blk * 32 + (l & 31), o.rows - 1);
    return o.p + (int64_t)r * o.ld + 8 * (l >> 5) + 4 * h;
  }
  const int c = min(row0 + blk * 32 + (l & 7) * 4, o.rows - 4);
  return o.p + (int64_t)(h * 8 + (l >> 3)) * o.ld + c;
}

// Lane l's fragment of block blk (8 consecutive k of its row) from a stage;
// KMAJOR is the operand's layout (a compile-time choice, so a stage's reads,
// splits and MFMAs form one block the scheduler can interleave).
template <bool KMAJOR>
__device__ __forceinline__ void direct_frag(const u32x4* st, int blk, int l, float (&x)[8]) {
  if constexpr(KMAJOR) {
    const floatx4 u = __builtin_bit_cast(floatx4, st[blk * 128 + l]);
    const floatx4 v = __builtin_bit_cast(floatx4, st[blk * 128 + 64 + l]);
#pragma unroll
    for(int e = 0; e < 4; ++e) {
      x[e] = u[e];
      x[4 + e] = v[e];
    }
  } else {
    const float* f = reinterpret_cast<const float*>(st + blk * 128);
#pragma unroll
    for(int e = 0; e < 8; ++e) x[e] = f[(8 * (l >> 5) + e) * 32 + (l & 31)];
  }
}

// Pair p (elements 2p, 2p+1) of split_chunk: the same operations, so a
// fragment split pair by pair equals split_chunk's.
template <int NHL>
__device__ __forceinline__ void split_pair(const float (&v)[8], int p, uint32_t& h, uint32_t& l) {
  h = cvt_pk_bf16(v[2 * p], v[2 * p + 1]);
  if(NHL == 2) {
    const float h0 = __builtin_bit_cast(float, h << 16);
    const float h1 = __builtin_bit_cast(float, h & 0xffff0000u);
    l = cvt_pk_bf16(v[2 * p] - h0, v[2 * p + 1] - h1);
  } else {
    l = 0;
  }
}

template <int BM, int BN, int WM, int WN, int NHL, int NB>
__device__ __forceinline__ void bf16_tile_direct(const DirectTileArgs& t, int tm, int tn, u32x4* lds) {
  using Cfg = BfDirectCfg<BM, BN, WM, WN, NB>;
  constexpr int NT = Cfg::NT, NW = Cfg::NW, FM = Cfg::FM, FN = Cfg::FN;
  constexpr int SA = Cfg::CH_A, SB = Cfg::CH_B;
  constexpr int PER = Cfg::PER_A + Cfg::PER_B;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  // The WM x WN sub-tiles that meet the problem go to the first waves, row by
  // row (a full tile: wave w takes sub-tile w, as everywhere else); the waves
  // past them still issue their share of every stage's loads but skip the
  // split, the MFMAs and the store.  A tile that covers a fraction of its
  // area (small and ragged problems) then keeps fewer waves per SIMD busy
  // instead of contracting padding.  Which wave computes a sub-tile does not
  // change its MFMA sequence, so results are bit-identical.
  const int act_m = min(Cfg::NWM, (t.ep.M - tm * BM + WM - 1) / WM);
  const int act_n = min(Cfg::NWN, (t.ep.N - tn * BN + WN - 1) / WN);
  const bool active = wid < act_m * act_n;
  const int wm = active ? wid / act_n : 0, wn = active ? wid % act_n : 0;

  floatx16 acc[FM][FN];
  zero_acc(acc);

  const int steps = t.K / 16;
  const int wave_base = __builtin_amdgcn_readfirstlane(tid - lane);
  // per DMA instruction: its source at k = 0; a stage advances k by 16
  const float* src_a[Cfg::PER_A];
  const float* src_b[Cfg::PER_B];
#pragma unroll
  for(int i = 0; i < Cfg::PER_A; ++i) src_a[i] = direct_src(t.a, tm * BM, wid + i * NW, lane);
#pragma unroll
  for(int i = 0; i < Cfg::PER_B; ++i) src_b[i] = direct_src(t.b, tn * BN, wid + i * NW, lane);
  const int64_t step_a = t.a.kmajor ? 16 : 16 * t.a.ld, step_b = t.b.kmajor ? 16 : 16 * t.b.ld;
  auto issue = [&](int s) __attribute__((always_inline)) {
    u32x4* sa = lds + (s % NB) * (SA + SB);
    u32x4* sb = sa + SA;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i)
      lds_dma16(reinterpret_cast<const u32x4*>(src_a[i] + s * step_a), sa + wave_base + i * NT);
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i)
      lds_dma16(reinterpret_cast<const u32x4*>(src_b[i] + s * step_b), sb + wave_base + i * NT);
  };
  // A stage, software-pipelined inside the wave (a wave issues in order, so
  // VALU work placed between MFMAs runs in their shadow): the B fragments and
  // A fragment 0 are split up front; then for each A fragment i its MFMAs,
  // each followed by one pair of A fragment i+1's split, while the LDS reads
  // of fragment i+2 are in flight.  sched_barrier(0) keeps that order.  Per
  // accumulator the products stay lo*hi, hi*lo, hi*hi, so results are those
  // of the unpipelined order.
  // A wave whose sub-tile the problem covers only in part (PARTIAL) also
  // skips, within it, the 32-row A blocks (i >= va) and 32-column B blocks
  // (j >= vb) that miss the problem: their LDS reads, splits and MFMAs.  That
  // copy of the loop runs only for such waves; a wholly covered wave runs the
  // loop without those tests.
  const int va = active ? min(FM, (t.ep.M - (tm * BM + wm * WM) + 31) / 32) : 0;
  const int vb = active ? min(FN, (t.ep.N - (tn * BN + wn * WN) + 31) / 32) : 0;
  const bool whole = va == FM && vb == FN;
  auto contract = [&](const u32x4* sa, const u32x4* sb, auto ak, auto bk, auto part) __attribute__((always_inline)) {
    constexpr bool AK = decltype(ak)::value, BKM = decltype(bk)::value, PARTIAL = decltype(part)::value;
    constexpr int MPF = NHL == 2 ? 3 * FN : FN;  // MFMAs per A fragment
    constexpr int PPM = (4 + MPF - 1) / MPF;     // split pairs after each MFMA
    u32x4 bh[FN], bl[FN];
    float xa[FM][8];
    u32x4 ah[FM], al[FM];
#pragma unroll
    for(int j = 0; j < FN; ++j) {
      if(!PARTIAL || j < vb) {
        float x[8];
        direct_frag<BKM>(sb, wn * FN + j, lane, x);
        split_chunk<NHL>(x, bh[j], bl[j]);
      }
    }
    direct_frag<AK>(sa, wm * FM, lane, xa[0]);
    if constexpr(FM > 1)
      if(!PARTIAL || 1 < va) direct_frag<AK>(sa, wm * FM + 1, lane, xa[1]);
    split_chunk<NHL>(xa[0], ah[0], al[0]);
    __builtin_amdgcn_sched_barrier(0);
    static_for<FM>([&](auto ii) __attribute__((always_inline)) {
      constexpr int i = decltype(ii)::value;
      if(PARTIAL && i >= va) return;
      if constexpr(i + 2 < FM)
        if(!PARTIAL || i + 2 < va) direct_frag<AK>(sa, wm * FM + i + 2, lane, xa[i + 2]);
      static_for<MPF>([&](auto mm) __attribute__((always_inline)) {
        constexpr int m = decltype(mm)::value;
        constexpr int prod = NHL == 2 ? m / FN : 2, j = m % FN;
        if(!PARTIAL || j < vb) {
          if constexpr(prod == 0)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(al[i]), as_bf16x8(bh[j]), acc[i][j], 0, 0, 0);
          else if constexpr(prod == 1)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(ah[i]), as_bf16x8(bl[j]), acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(ah[i]), as_bf16x8(bh[j]), acc[i][j], 0, 0, 0);
        }
        if constexpr(i + 1 < FM) {
          if(!PARTIAL || i + 1 < va) {
#pragma unroll
            for(int q = 0; q < PPM; ++q) {
              const int p = m * PPM + q;
              if(p < 4) {
                uint32_t hw, lw;
                split_pair<NHL>(xa[i + 1], p, hw, lw);
                ah[i + 1][p] = hw;
                al[i + 1][p] = lw;
              }
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      });
    });
  };

  auto k_loop = [&](auto ak, auto bk, auto part) __attribute__((always_inline)) {
#pragma unroll
    for(int p = 0; p < NB - 1; ++p)
      if(p < steps) issue(p);
    for(int s = 0; s < steps; ++s) {
      const int after = steps - 1 - s;
      if(after >= NB - 2) vm_wait<(NB - 2) * PER>();
      else if(NB > 3 && after == 1) vm_wait<PER>();
      else lds_dma_wait();
      __syncthreads();
      if(s + NB - 1 < steps) issue(s + NB - 1);
      const u32x4* sa = lds + (s % NB) * (SA + SB);
      if(active) contract(sa, sa + SA, ak, bk, part);
    }
  };
  // one instance of the loop per pair of operand layouts and coverage (every
  // instance passes the same barriers, so the waves of a workgroup may run
  // different ones)
  using T = std::true_type;
  using F = std::false_type;
  auto by_cover = [&](auto ak, auto bk) __attribute__((always_inline)) {
    if(whole) k_loop(ak, bk, F{});
    else k_loop(ak, bk, T{});
  };
  if(t.a.kmajor) {
    if(t.b.kmajor) by_cover(T{}, T{});
    else by_cover(T{}, F{});
  } else {
    if(t.b.kmajor) by_cover(F{}, T{});
    else by_cover(F{}, F{});
  }

  if(active) static_for<FM * FN>([&](auto ij) __attribute__((always_inline)) {
    constexpr int i = decltype(ij)::value / FN, j = decltype(ij)::value % FN;
    store_block32(t.ep, acc[i][j], tm * BM + wm * WM + i * 32, tn * BN + wn * WN + j * 32, lane);
  });
}

// Host: whether an operand can be read by the direct tile.
inline bool direct_operand_ok(const float* p, int64_t ld, int rows, bool kmajor) {
  if((reinterpret_cast<uintptr_t>(p) & 15u) != 0 || ld % 4 != 0) return false;
  return kmajor || rows % 4 == 0;
}

}  // namespace mfx
