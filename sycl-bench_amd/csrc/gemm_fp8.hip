// Block-scaled FP8 (OCP e4m3) contractions on v_mfma_scale_f32_32x32x64_f8f6f4.
//
// Quantisation (one workgroup per 128 x 128 block of an operand): the block's
// largest magnitude fixes a power-of-two scale 2^e (the smallest e with
// amax <= 448 * 2^e), every element becomes q = e4m3_rne(x * 2^-e).  Powers of
// two make the scale an E8M0 byte that the MFMA applies itself, per operand,
// so products need no rescaling pass.  Two-term mode quantises the residual
// x - q1 * 2^e1 the same way (q2, e2) and contracts q1*q1 + q1*q2 + q2*q1.
//
// Packed layout (A role: rows = M; B role: rows = N), fragment order of the
// 32x32x64 MFMA: P[blk32][kb64][half][lane][16 bytes], lane l holding row
// blk32*32 + l%32 and k = kb64*64 + 32*(l/32) + 16*half + 0..15.  Rows and k
// are zero-padded to the tile and to 128.  Exponents: E[row/128][k/128].
//
// Contraction tiles: 256x256 (four waves of 128x128, or eight of 128x64) or
// 128x128 (waves of 64x64);
// each wave's rows and columns lie inside one 128-block, so its A and B
// scales are one byte each per stage.  Stage = 128 k (single) or 64 k
// (two-term), double-buffered in LDS and filled by direct global-to-LDS
// loads, 16-byte chunks read conflict-free.
#include "mfx_internal.h"

namespace mfx {

int fp8_config_count();
bool fp8_config_info(int c, int* bm, int* bn, int* bk, int* waves);

namespace {

typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Quantisation
// ---------------------------------------------------------------------------
struct QuantParams {
  const float* src;
  int64_t ld;
  int rows, K;
  int kmajor;       // element (r, k) at src[r*ld + k], else src[k*ld + r]
  int vec;          // kmajor with 16-byte aligned rows: 128-bit loads
  int kb128s;       // padded K / 128
  int two_term;
  u32x4* q1;        // packed, fragment order
  u32x4* q2;
  int* e1;          // [rows_pad/128][kb128s]
  int* e2;
  // plain row-major output (mfx_quantize_fp8_block), used when q1 == nullptr
  uint8_t* plain;
  int64_t plain_ld;
  int block;        // 128 for the contraction; 32/64/128 for the utility
};

__device__ __forceinline__ float block_max(float v, float* sh) {
  for(int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  float m = sh[0];
  for(int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmaxf(m, sh[i]);
  return m;
}

// e4m3 codes of four floats, packed little-endian into one word, by the
// hardware conversion (v_cvt_pk_fp8_f32: OCP E4M3FN, round to nearest even).
// Inputs are scaled so |x| <= 448 (the block scale guarantees it), so no
// saturation is needed.
__device__ __forceinline__ uint32_t e4m3_pack4(float a, float b, float c, float d) {
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
  return (uint32_t)w;
}

// Value of byte `j` (0..3) of a word of e4m3 codes (v_cvt_f32_fp8).
template <int J>
__device__ __forceinline__ float e4m3_byte(uint32_t w) {
  return __builtin_amdgcn_cvt_f32_fp8((int)w, J);
}

// One workgroup per (row block, k block) of `block` x `block` elements;
// thread t handles chunks of 16 consecutive k of one row.  k-contiguous
// sources put consecutive threads along k (64-byte vector loads per
// thread); row-contiguous ones put them along the rows (coalesced across
// lanes).
// Row pitch (bytes) of the LDS code tile: 128 codes + 16, so 16-byte reads
// of one column by consecutive rows spread over the banks.
constexpr int QPITCH = 144;

__global__ __launch_bounds__(256) void quantize_fp8_kernel(const QuantParams p) {
  __shared__ float sh[4];
  __shared__ __align__(16) unsigned char qtile[2][128 * QPITCH];
  // Neighbouring workgroups take neighbouring blocks along the source's
  // contiguous dimension (k if k-major, rows if not), so they read
  // neighbouring segments of the same memory rows.
  const int rb = p.kmajor ? blockIdx.y : blockIdx.x, kb = p.kmajor ? blockIdx.x : blockIdx.y;
  const int B = p.block;
  const int per_row = B / 16;
  const int chunks = B * per_row;  // <= 1024
  float v[4][16];
  float amax = 0.0f;
#pragma unroll
  for(int c = 0; c < 4; ++c) {
    const int id = threadIdx.x + c * 256;
    const int rr = p.kmajor ? id / per_row : id % B, kk = p.kmajor ? id % per_row : id / B;
    const int r = rb * B + rr, k0 = kb * B + kk * 16;
    const bool live = id < chunks && r < p.rows;
    if(p.kmajor) {
      const float* src = p.src + (int64_t)r * p.ld + k0;
      if(live && p.vec && k0 + 16 <= p.K) {
#pragma unroll
        for(int q = 0; q < 4; ++q) {
          const float4 x = reinterpret_cast<const float4*>(src)[q];
          v[c][4 * q] = x.x;
          v[c][4 * q + 1] = x.y;
          v[c][4 * q + 2] = x.z;
          v[c][4 * q + 3] = x.w;
        }
      } else {
#pragma unroll
        for(int e = 0; e < 16; ++e) v[c][e] = (live && k0 + e < p.K) ? src[e] : 0.0f;
      }
    } else {
#pragma unroll
      for(int e = 0; e < 16; ++e) v[c][e] = (live && k0 + e < p.K) ? p.src[(int64_t)(k0 + e) * p.ld + r] : 0.0f;
    }
#pragma unroll
    for(int e = 0; e < 16; ++e) amax = fmaxf(amax, fabsf(v[c][e]));
  }
  const int ex1 = fp8_block_exponent(block_max(amax, sh));
  const float inv1 = ldexpf(1.0f, -ex1), s1 = ldexpf(1.0f, ex1);
  uint32_t w1[4][4];
  float amax2 = 0.0f;
#pragma unroll
  for(int c = 0; c < 4; ++c)
#pragma unroll
    for(int j = 0; j < 4; ++j) {
      float* x = &v[c][4 * j];
      w1[c][j] = e4m3_pack4(x[0] * inv1, x[1] * inv1, x[2] * inv1, x[3] * inv1);
      if(p.two_term) {
        x[0] -= e4m3_byte<0>(w1[c][j]) * s1;
        x[1] -= e4m3_byte<1>(w1[c][j]) * s1;
        x[2] -= e4m3_byte<2>(w1[c][j]) * s1;
        x[3] -= e4m3_byte<3>(w1[c][j]) * s1;
        amax2 = fmaxf(amax2, fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3]))));
      }
    }
  int ex2 = 0;
  float inv2 = 1.0f;
  if(p.two_term) {
    ex2 = fp8_block_exponent(block_max(amax2, sh));
    inv2 = ldexpf(1.0f, -ex2);
  }
  if(threadIdx.x == 0) {
    p.e1[rb * p.kb128s + kb] = ex1;
    if(p.two_term) p.e2[rb * p.kb128s + kb] = ex2;
  }
#pragma unroll
  for(int c = 0; c < 4; ++c) {
    const int id = threadIdx.x + c * 256;
    if(id >= chunks) continue;
    const int rr = p.kmajor ? id / per_row : id % B, kk = p.kmajor ? id % per_row : id / B;
    const int r = rb * B + rr, k0 = kb * B + kk * 16;
    uint32_t w2[4] = {0, 0, 0, 0};
    if(p.two_term)
#pragma unroll
      for(int j = 0; j < 4; ++j)
        w2[j] = e4m3_pack4(v[c][4 * j] * inv2, v[c][4 * j + 1] * inv2, v[c][4 * j + 2] * inv2, v[c][4 * j + 3] * inv2);
    if(p.q1) {
      // codes of (rr, kk*16 .. kk*16+15) into the block's LDS tile
      *reinterpret_cast<u32x4*>(&qtile[0][rr * QPITCH + kk * 16]) = (u32x4){w1[c][0], w1[c][1], w1[c][2], w1[c][3]};
      if(p.two_term) *reinterpret_cast<u32x4*>(&qtile[1][rr * QPITCH + kk * 16]) = (u32x4){w2[0], w2[1], w2[2], w2[3]};
    } else if(r < p.rows) {
      for(int e = 0; e < 16 && k0 + e < p.K; ++e)
        p.plain[(int64_t)r * p.plain_ld + k0 + e] = (uint8_t)(w1[c][e / 4] >> (8 * (e % 4)));
    }
  }
  if(p.q1) {
    // The block's 1024 fragment-order chunks (block = 128 here), consecutive
    // threads on consecutive chunks: chunk (blk, kb64, half, lane) holds the
    // codes of row blk*32 + lane%32, k = kb64*64 + 32*(lane/32) + 16*half + 0..15.
    __syncthreads();
    const int kb64s = p.kb128s * 2;
#pragma unroll
    for(int c = 0; c < 4; ++c) {
      const int ch = threadIdx.x + c * 256;
      const int lane = ch & 63, half = (ch >> 6) & 1, kb64l = (ch >> 7) & 1, blkl = ch >> 8;
      const int row = blkl * 32 + (lane & 31), col = kb64l * 64 + 32 * (lane >> 5) + 16 * half;
      const int64_t at = ((((int64_t)(rb * 4 + blkl)) * kb64s + kb * 2 + kb64l) * 2 + half) * 64 + lane;
      p.q1[at] = *reinterpret_cast<const u32x4*>(&qtile[0][row * QPITCH + col]);
      if(p.two_term) p.q2[at] = *reinterpret_cast<const u32x4*>(&qtile[1][row * QPITCH + col]);
    }
  }
}

// ---------------------------------------------------------------------------
// Contraction
// ---------------------------------------------------------------------------
struct Fp8TileArgs {
  const u32x4* A1;
  const u32x4* A2;
  const u32x4* B1;
  const u32x4* B2;
  const int* eA1;
  const int* eA2;
  const int* eB1;
  const int* eB2;
  int kb128s;  // padded K / 128
  int tiles_m, tiles_n, group;
  Epilogue ep;
};

template <int BM, int BN, int WM, int WN, int T>
struct Fp8Cfg {
  static constexpr int NWN = BN / WN, NT = (BM / WM) * (BN / WN) * 64;
  static constexpr int FM = WM / 32, FN = WN / 32;
  static constexpr int KS = T == 1 ? 2 : 1;                   // 64-wide k steps per stage
  static constexpr int RUN = KS * 2 * 64;                     // chunks of one 32-row block per stage and term
  static constexpr int CH_A = BM / 32 * RUN * T, CH_B = BN / 32 * RUN * T;
  static constexpr int PER_A = CH_A / NT, PER_B = CH_B / NT;
  static_assert(CH_A % NT == 0 && CH_B % NT == 0, "stage must split evenly");
  static_assert(WM <= 128 && 128 % WM == 0 && WN <= 128 && 128 % WN == 0, "a wave must sit in one scale block");
};

__device__ __forceinline__ i32x8 join(u32x4 lo, u32x4 hi) {
  i32x8 r;
  r[0] = lo.x; r[1] = lo.y; r[2] = lo.z; r[3] = lo.w;
  r[4] = hi.x; r[5] = hi.y; r[6] = hi.z; r[7] = hi.w;
  return r;
}

template <int BM, int BN, int WM, int WN, int T>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_fp8_kernel(const Fp8TileArgs t) {
  using Cfg = Fp8Cfg<BM, BN, WM, WN, T>;
  constexpr int NT = Cfg::NT, FM = Cfg::FM, FN = Cfg::FN, KS = Cfg::KS, RUN = Cfg::RUN;
  constexpr int SA = Cfg::CH_A, SB = Cfg::CH_B;
  __shared__ u32x4 lds[2 * (SA + SB)];

  int tm, tn;
  grouped_tile(xcd_remap(blockIdx.x, gridDim.x), t.tiles_m, t.tiles_n, t.group, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid / Cfg::NWN, wn = wid % Cfg::NWN;
  const int arow128 = (tm * BM + wm * WM) / 128, bcol128 = (tn * BN + wn * WN) / 128;

  floatx16 acc[FM][FN];
#pragma unroll
  for(int i = 0; i < FM; ++i)
#pragma unroll
    for(int j = 0; j < FN; ++j)
#pragma unroll
      for(int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  const int kb64s = t.kb128s * 2;
  const int steps = T == 1 ? t.kb128s : kb64s;
  const int64_t blk_stride = (int64_t)kb64s * 2 * 64;

  // Global -> LDS without registers (global_load_lds_dwordx4): for one wave
  // and one i the 64 lanes' chunks are contiguous both in the packed operand
  // (RUN and SA / T are multiples of 64) and in the LDS stage.
  const int wave_base = tid - lane;
  auto issue = [&](int s, int buf) __attribute__((always_inline)) {
    const int64_t koff = (int64_t)s * RUN;
    u32x4* sa = lds + buf * (SA + SB);
    u32x4* sb = sa + SA;
#pragma unroll
    for(int i = 0; i < Cfg::PER_A; ++i) {
      const int c = tid + i * NT;
      const int term = c / (SA / T), cc = c % (SA / T);
      const int blk = cc / RUN, run = cc - blk * RUN;
      const u32x4* src = term ? t.A2 : t.A1;
      lds_dma16(src + (int64_t)(tm * (BM / 32) + blk) * blk_stride + koff + run, sa + wave_base + i * NT);
    }
#pragma unroll
    for(int i = 0; i < Cfg::PER_B; ++i) {
      const int c = tid + i * NT;
      const int term = c / (SB / T), cc = c % (SB / T);
      const int blk = cc / RUN, run = cc - blk * RUN;
      const u32x4* src = term ? t.B2 : t.B1;
      lds_dma16(src + (int64_t)(tn * (BN / 32) + blk) * blk_stride + koff + run, sb + wave_base + i * NT);
    }
  };
  // E8M0 scale bytes (biased exponents) of this wave's A rows and B columns
  // for stage s, fetched one stage ahead.
  struct Scales {
    int a1, b1, a2, b2;
  };
  auto scales = [&](int s) __attribute__((always_inline)) {
    const int kb128 = T == 1 ? s : s / 2;
    Scales r;
    r.a1 = t.eA1[arow128 * t.kb128s + kb128] + 127;
    r.b1 = t.eB1[bcol128 * t.kb128s + kb128] + 127;
    r.a2 = T == 2 ? t.eA2[arow128 * t.kb128s + kb128] + 127 : 0;
    r.b2 = T == 2 ? t.eB2[bcol128 * t.kb128s + kb128] + 127 : 0;
    return r;
  };
  auto compute = [&](int cur, const Scales& sc) __attribute__((always_inline)) {
    const u32x4* sa = lds + cur * (SA + SB);
    const u32x4* sb = sa + SA;
#pragma unroll
    for(int ks = 0; ks < KS; ++ks) {
      i32x8 a1[FM], a2[FM], b1[FN], b2[FN];
#pragma unroll
      for(int i = 0; i < FM; ++i) {
        const int base = (wm * FM + i) * RUN + ks * 128 + lane;
        a1[i] = join(sa[base], sa[base + 64]);
        if(T == 2) a2[i] = join(sa[SA / 2 + base], sa[SA / 2 + base + 64]);
      }
#pragma unroll
      for(int j = 0; j < FN; ++j) {
        const int base = (wn * FN + j) * RUN + ks * 128 + lane;
        b1[j] = join(sb[base], sb[base + 64]);
        if(T == 2) b2[j] = join(sb[SB / 2 + base], sb[SB / 2 + base + 64]);
      }
      if(T == 2) {
#pragma unroll
        for(int i = 0; i < FM; ++i)
#pragma unroll
          for(int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a2[i], b1[j], acc[i][j], 0, 0, 0, sc.a2, 0, sc.b1);
#pragma unroll
        for(int i = 0; i < FM; ++i)
#pragma unroll
          for(int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1[i], b2[j], acc[i][j], 0, 0, 0, sc.a1, 0, sc.b2);
      }
#pragma unroll
      for(int i = 0; i < FM; ++i)
#pragma unroll
        for(int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1[i], b1[j], acc[i][j], 0, 0, 0, sc.a1, 0, sc.b1);
    }
  };

  // Stage s+1 streams into the other buffer (and its scales into registers)
  // while stage s is contracted; one barrier per stage.
  Scales cur{127, 127, 127, 127};
  if(steps > 0) {
    issue(0, 0);
    cur = scales(0);
  }
  lds_dma_wait();
  __syncthreads();
  for(int s = 0; s + 1 < steps; ++s) {
    issue(s + 1, (s + 1) & 1);
    const Scales next = scales(s + 1);
    compute(s & 1, cur);
    cur = next;
    lds_dma_wait();
    __syncthreads();
  }
  if(steps > 0) compute((steps - 1) & 1, cur);

  static_for<FM * FN>([&](auto ij) __attribute__((always_inline)) {
    constexpr int i = decltype(ij)::value / FN, j = decltype(ij)::value % FN;
    store_block32(t.ep, acc[i][j], tm * BM + wm * WM + i * 32, tn * BN + wn * WN + j * 32, lane);
  });
}

struct Fp8Config {
  int bm, bn, waves;  // waves = 8: 128x64 per wave, two waves per SIMD
};
constexpr Fp8Config kFp8Configs[] = {{256, 256, 4}, {128, 128, 4}, {256, 256, 8}};
constexpr int kFp8ConfigCount = 3;

struct Fp8Plan {
  int64_t rows_pad_a, rows_pad_b, kb128s;
  size_t q_a, q_b, e_a, e_b;  // bytes of one term
};

Fp8Plan fp8_plan(const DenseOp& op, int config) {
  const Fp8Config& c = kFp8Configs[config];
  Fp8Plan pl;
  pl.rows_pad_a = round_up(op.M, std::max(c.bm, 128));
  pl.rows_pad_b = round_up(op.N, std::max(c.bn, 128));
  pl.kb128s = round_up(std::max(op.K, 1), 128) / 128;
  pl.q_a = (size_t)pl.rows_pad_a * pl.kb128s * 128;
  pl.q_b = (size_t)pl.rows_pad_b * pl.kb128s * 128;
  pl.e_a = (size_t)(pl.rows_pad_a / 128) * pl.kb128s * 4;
  pl.e_b = (size_t)(pl.rows_pad_b / 128) * pl.kb128s * 4;
  return pl;
}

mfx_status quantize(Handle* h, const float* src, int64_t ld, int rows, int K, bool kmajor, int64_t rows_pad,
    int64_t kb128s, bool two, u32x4* q1, u32x4* q2, int* e1, int* e2, hipStream_t s) {
  QuantParams p{};
  p.src = src;
  p.ld = ld;
  p.rows = rows;
  p.K = K;
  p.kmajor = kmajor;
  p.vec = kmajor && (reinterpret_cast<uintptr_t>(src) & 15u) == 0 && ld % 4 == 0;
  p.kb128s = (int)kb128s;
  p.two_term = two;
  p.q1 = q1;
  p.q2 = q2;
  p.e1 = e1;
  p.e2 = e2;
  p.block = 128;
  const unsigned rbs = (unsigned)(rows_pad / 128);
  const dim3 grid = kmajor ? dim3((unsigned)kb128s, rbs) : dim3(rbs, (unsigned)kb128s);
  return timed_launch(h, two ? "quantize_fp8_block2" : "quantize_fp8_block", s,
      [&] { quantize_fp8_kernel<<<grid, 256, 0, s>>>(p); });
}

}  // namespace

int fp8_config_count() { return kFp8ConfigCount; }
bool fp8_config_info(int c, int* bm, int* bn, int* bk, int* waves) {
  if(c < 0 || c >= kFp8ConfigCount) return false;
  *bm = kFp8Configs[c].bm;
  *bn = kFp8Configs[c].bn;
  *bk = 128;
  *waves = kFp8Configs[c].waves;
  return true;
}

size_t dense_fp8_workspace(const DenseOp& op, bool two_term, int config) {
  if(config < 0 || config >= kFp8ConfigCount) return 0;
  const Fp8Plan pl = fp8_plan(op, config);
  const int terms = two_term ? 2 : 1;
  return terms * (round_up(pl.q_a, 256) + round_up(pl.q_b, 256) + round_up(pl.e_a, 256) + round_up(pl.e_b, 256));
}

mfx_status launch_dense_fp8(Handle* h, const DenseOp& op, bool two_term, int config, hipStream_t s) {
  if(config < 0 || config >= kFp8ConfigCount) return invalid("FP8 tile configuration out of range");
  if(op.nprod != 1 || op.tiles != TILES_FULL) return unsupported("FP8 is built for plain gemm (not syrk / syr2k)");
  if(op.M == 0 || op.N == 0) return MFX_SUCCESS;
  const Fp8Plan pl = fp8_plan(op, config);
  void* ws = nullptr;
  MFX_CHECK(workspace(h, dense_fp8_workspace(op, two_term, config), &ws));
  char* w = static_cast<char*>(ws);
  auto take = [&](size_t bytes) {
    char* p = w;
    w += round_up(bytes, 256);
    return p;
  };
  u32x4* qa1 = reinterpret_cast<u32x4*>(take(pl.q_a));
  u32x4* qb1 = reinterpret_cast<u32x4*>(take(pl.q_b));
  int* ea1 = reinterpret_cast<int*>(take(pl.e_a));
  int* eb1 = reinterpret_cast<int*>(take(pl.e_b));
  u32x4 *qa2 = nullptr, *qb2 = nullptr;
  int *ea2 = nullptr, *eb2 = nullptr;
  if(two_term) {
    qa2 = reinterpret_cast<u32x4*>(take(pl.q_a));
    qb2 = reinterpret_cast<u32x4*>(take(pl.q_b));
    ea2 = reinterpret_cast<int*>(take(pl.e_a));
    eb2 = reinterpret_cast<int*>(take(pl.e_b));
  }
  MFX_CHECK(quantize(h, op.A, op.lda, op.M, op.K, op.a_kmajor, pl.rows_pad_a, pl.kb128s, two_term, qa1, qa2, ea1, ea2, s));
  MFX_CHECK(quantize(h, op.B, op.ldb, op.N, op.K, op.b_kmajor, pl.rows_pad_b, pl.kb128s, two_term, qb1, qb2, eb1, eb2, s));
  Fp8TileArgs t;
  t.A1 = qa1;
  t.A2 = qa2 ? qa2 : qa1;
  t.B1 = qb1;
  t.B2 = qb2 ? qb2 : qb1;
  t.eA1 = ea1;
  t.eA2 = ea2 ? ea2 : ea1;
  t.eB1 = eb1;
  t.eB2 = eb2 ? eb2 : eb1;
  t.kb128s = (int)pl.kb128s;
  t.ep = op.ep;
  t.group = 8;
  const Fp8Config& c = kFp8Configs[config];
  t.tiles_m = (int)ceil_div(op.M, c.bm);
  t.tiles_n = (int)ceil_div(op.N, c.bn);
  const int grid = t.tiles_m * t.tiles_n;
  const char* name = two_term ? "gemm_fp8_block2_mfma" : "gemm_fp8_block_mfma";
  if(config == 2) {
    if(two_term)
      return timed_launch(h, name, s, [&] { gemm_fp8_kernel<256, 256, 128, 64, 2><<<grid, 512, 0, s>>>(t); });
    return timed_launch(h, name, s, [&] { gemm_fp8_kernel<256, 256, 128, 64, 1><<<grid, 512, 0, s>>>(t); });
  }
  if(config == 0) {
    if(two_term)
      return timed_launch(h, name, s, [&] { gemm_fp8_kernel<256, 256, 128, 128, 2><<<grid, 256, 0, s>>>(t); });
    return timed_launch(h, name, s, [&] { gemm_fp8_kernel<256, 256, 128, 128, 1><<<grid, 256, 0, s>>>(t); });
  }
  if(two_term) return timed_launch(h, name, s, [&] { gemm_fp8_kernel<128, 128, 64, 64, 2><<<grid, 256, 0, s>>>(t); });
  return timed_launch(h, name, s, [&] { gemm_fp8_kernel<128, 128, 64, 64, 1><<<grid, 256, 0, s>>>(t); });
}

mfx_status launch_quantize_fp8(Handle* h, int64_t rows, int64_t cols, const float* x, int64_t ldx, int block,
    uint8_t* q, int32_t* exps, hipStream_t s) {
  if(rows == 0 || cols == 0) return MFX_SUCCESS;
  QuantParams p{};
  p.src = x;
  p.ld = ldx;
  p.rows = (int)rows;
  p.K = (int)cols;
  p.kmajor = 1;
  p.vec = (reinterpret_cast<uintptr_t>(x) & 15u) == 0 && ldx % 4 == 0;
  p.kb128s = (int)ceil_div(cols, block);
  p.two_term = 0;
  p.e1 = exps;
  p.plain = q;
  p.plain_ld = cols;
  p.block = block;
  dim3 grid((unsigned)ceil_div(cols, block), (unsigned)ceil_div(rows, block));
  return timed_launch(h, "quantize_fp8_block_plain", s, [&] { quantize_fp8_kernel<<<grid, 256, 0, s>>>(p); });
}

}  // namespace mfx
