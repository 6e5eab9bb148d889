// Body of the fused training step on BF16 MFMA (gfx950), shared by the
// kernels of kernels_fused.hip (the training step, several sets per launch)
// and kernels_fused_pair.hip (the forward / backward pair): gather + decoder
// forward + loss + hand-written backward + latent-gradient scatter +
// per-workgroup decoder gradient accumulation, in one persistent kernel. Replaces, per step, the
// reference's draw_features, draw_targets, 9 cuBLAS SGEMMs, 4 HardGELU
// launches, Saxpy/Sdot/scale and accumulate_grid_gradients
// (libopenntc.cpp:621-709, ntc_kernel.cu:75-373).
//
// Work split: a workgroup (4 waves) takes passes of 128 texels; each wave
// owns 32 texels (two 16-texel MFMA column tiles) and runs the whole
// forward/backward chain for them in registers with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation):
//   * layer outputs stay in the MFMA accumulator layout (lane l holds rows
//     4(l/16)+i of column l%16) and feed the next layer's B operand directly:
//     the next layer's weight columns are stored permuted so that B-operand
//     position (l/16, e) of K-chunk j means hidden unit
//     kperm = 32j + (e<4 ? 4(l/16)+e : 16+4(l/16)+e-4);
//   * all weight tables (W0, W1, Wout and the transposes the backward needs,
//     BF16, permuted) sit in LDS for the whole kernel;
//   * decoder weight gradients need the texel index as the MFMA K dimension,
//     so activations are staged through LDS in [unit][texel] layout and the
//     four waves each accumulate a quarter of dW in registers across passes;
//     the partial sums go to HBM once per workgroup and are reduced in order
//     by ntc_decoder_update (deterministic);
//   * latent gradients: dX of the grid slots goes to LDS per texel; each
//     (cell, channel) of the pass's footprint box then sums its texels'
//     weighted dX in a fixed order and adds the sum to the grid's gradient
//     with one global float atomic (the reference adds every texel's
//     contribution with an atomic, ntc_kernel.cu:261-278).
// Quantisation noise is the same Philox stream as the plain path.
#pragma once

#include <cstring>

#include "ntc_device.h"
#include "ntc_host.h"

using namespace ntc;

namespace
{
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 frag_lds(const uint16_t* p)
{
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4i16 lds_v4i16;

// ds_read_b64_tr_b16: within each 16-lane group, lane i receives element
// (i & 3) of the 8 bytes that lanes 4k + (i >> 2) address, for k = 0..3. When
// lane 4k + q addresses four consecutive BF16 of row R(k) at column block q of
// a row-major table, lane i therefore gets column i of rows R(0..3): four
// elements of a transposed fragment.
__device__ __forceinline__ uint2 tr_read(const uint16_t* p)
{
  const v4i16 v = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4i16*)p);
  return __builtin_bit_cast(uint2, v);
}

// A fragment of a transposed table from two transposing reads: elements 0..3
// from rows R(0..3) (address p0), 4..7 from rows R(4..7) (address p1).
__device__ __forceinline__ bf16x8 frag_tr(const uint16_t* p0, const uint16_t* p1)
{
  const uint2 a = tr_read(p0), b = tr_read(p1);
  return __builtin_bit_cast(bf16x8, make_uint4(a.x, a.y, b.x, b.y));
}

// Elements 0..3 from one transposing read, 4..7 zero.
__device__ __forceinline__ bf16x8 frag_tr_half(const uint16_t* p0)
{
  const uint2 a = tr_read(p0);
  return __builtin_bit_cast(bf16x8, make_uint4(a.x, a.y, 0u, 0u));
}

// body(i, load(i)) for i = 0 .. N-1, each fragment loaded from LDS RING uses
// ahead: with one wave per SIMD nothing else hides a read's latency, and the
// compiler left alone reads each fragment just before its MFMA.
// The layers walk their K chunks outermost (i = j * MT + r): consecutive MFMAs
// write different accumulators, and each accumulator still sums its K chunks
// in order.
template <int N, int RING, class Load, class Body>
__device__ __forceinline__ void lds_ring_ld(Load load, Body body)
{
  constexpr int R = RING < N ? RING : N;
  bf16x8 ring[R];
#pragma unroll
  for (int q = 0; q < R; ++q) ring[q] = load(q);
#pragma unroll
  for (int i = 0; i < N; ++i)
  {
    const bf16x8 a = ring[i % R];
    if (i + R < N) ring[i % R] = load(i + R);
    body(i, a);
  }
}

// lds_ring_ld of plain 16-byte fragment reads at addr(i).
template <int N, int RING, class Addr, class Body>
__device__ __forceinline__ void lds_ring(Addr addr, Body body)
{
  lds_ring_ld<N, RING>([&](int i) { return frag_lds(addr(i)); }, body);
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// lo, hi to BF16 (round to nearest even), packed lo | hi << 16: one
// v_cvt_pk_bf16_f32, the bits f2bf gives for every finite value.
__device__ __forceinline__ uint32_t pk_bf16(float lo, float hi)
{
  const f32x2_t v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

__device__ __forceinline__ uint16_t bf16_dev(float f) { return (uint16_t)pk_bf16(f, 0.0f); }

__device__ __forceinline__ bf16x8 frag_pack(const float* v)
{
  uint4 u;
  u.x = pk_bf16(v[0], v[1]);
  u.y = pk_bf16(v[2], v[3]);
  u.z = pk_bf16(v[4], v[5]);
  u.w = pk_bf16(v[6], v[7]);
  return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ uint16_t frag_elem(const bf16x8& f, int e)
{
  const uint4 u = __builtin_bit_cast(uint4, f);
  const uint32_t w = e < 2 ? u.x : (e < 4 ? u.y : (e < 6 ? u.z : u.w));
  return (uint16_t)((e & 1) ? (w >> 16) : (w & 0xFFFFu));
}

typedef short v2i16 __attribute__((ext_vector_type(2)));
typedef unsigned short v2u16 __attribute__((ext_vector_type(2)));

// ReLU of two packed BF16 values (one v_pk_max_i16): read as signed 16-bit
// integers the negative values (sign bit set, -0 included) lie below zero,
// the others keep their order, so max(v, 0) gives the BF16 ReLU. Rounding
// to BF16 and the ReLU commute, so relu_pk(pk_bf16(a, b)) has the bits of
// the ReLU applied in FP32 and then rounded.
__device__ __forceinline__ uint32_t relu_pk(uint32_t w)
{
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(v2i16, w), v2i16{0, 0}));
}

// ReLU backward of two packed BF16 values: d where the paired activation a
// (a ReLU output, so +0 or positive) is nonzero, else +0. The BF16
// activation is nonzero exactly when the pre-activation is positive, bar
// positive values below 2^-134 that round to zero, so the FP32
// pre-activations need not stay live through the backward. Masking after
// the rounding gives the bits of masking in FP32 first.
__device__ __forceinline__ uint32_t relu_bwd_pk(uint32_t d, uint32_t a)
{
  // per half: 0 - a has its sign bit set exactly when 0 < a <= 0x7FFF; the
  // arithmetic shift spreads it into an all-ones or all-zeros mask
  const v2i16 m = (v2i16{0, 0} - __builtin_bit_cast(v2i16, a)) >> 15;
  return d & __builtin_bit_cast(uint32_t, m);
}

// b / 255.0f for a byte b, correctly rounded, without a division: q = b r
// with r = fl(1/255), then one FMA correction; equal to the IEEE quotient for
// every b in 0..255 (checked exhaustively in exact arithmetic).
__device__ __forceinline__ float byte_over_255(uint32_t b)
{
  constexpr float r = 1.0f / 255.0f;
  const float x = (float)b;
  const float q = x * r;
  const float e = fmaf(-q, 255.0f, x);
  return fmaf(e, r, q);
}

// Two accumulator tiles (hidden units 4 (l/16) + i of tiles 2 j and 2 j + 1)
// as one B fragment of the next product (kperm order), after the activation.
template <int ACT>
__device__ __forceinline__ bf16x8 act_frag(const f32x4& lo, const f32x4& hi)
{
  if constexpr (ACT == NTC_ACT_RELU)
  {
    uint4 u;
    u.x = relu_pk(pk_bf16(lo[0], lo[1]));
    u.y = relu_pk(pk_bf16(lo[2], lo[3]));
    u.z = relu_pk(pk_bf16(hi[0], hi[1]));
    u.w = relu_pk(pk_bf16(hi[2], hi[3]));
    return __builtin_bit_cast(bf16x8, u);
  }
  else
  {
    float v[8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      v[i] = act_fwd<float>(ACT, lo[i]);
      v[4 + i] = act_fwd<float>(ACT, hi[i]);
    }
    return frag_pack(v);
  }
}

// The backward of the activation for two accumulator tiles of upstream
// gradients (dA) as one B fragment: ReLU masks the packed BF16 gradient with
// the activation fragment af of the same positions; other activations use
// the FP32 pre-activations zlo / zhi.
template <int ACT>
__device__ __forceinline__ bf16x8 act_bwd_frag(const f32x4& dlo, const f32x4& dhi, const bf16x8& af,
                                               const f32x4& zlo, const f32x4& zhi)
{
  if constexpr (ACT == NTC_ACT_RELU)
  {
    const uint4 a = __builtin_bit_cast(uint4, af);
    uint4 u;
    u.x = relu_bwd_pk(pk_bf16(dlo[0], dlo[1]), a.x);
    u.y = relu_bwd_pk(pk_bf16(dlo[2], dlo[3]), a.y);
    u.z = relu_bwd_pk(pk_bf16(dhi[0], dhi[1]), a.z);
    u.w = relu_bwd_pk(pk_bf16(dhi[2], dhi[3]), a.w);
    return __builtin_bit_cast(bf16x8, u);
  }
  else
  {
    float v[8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      v[i] = act_bwd<float>(ACT, zlo[i], dlo[i]);
      v[4 + i] = act_bwd<float>(ACT, zhi[i], dhi[i]);
    }
    return frag_pack(v);
  }
}

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Hidden unit held at B-operand position (g, e) of K-chunk j.
__host__ __device__ __forceinline__ int kperm(int j, int g, int e)
{
  return 32 * j + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

__host__ __device__ __forceinline__ int kperm_pos(int pos)
{
  const int j = pos >> 5, r = pos & 31;
  return kperm(j, r >> 3, r & 7);
}

// Inverse of kperm_pos: the permuted column that holds index k.
__host__ __device__ __forceinline__ int kperm_col(int k)
{
  const int r = k & 31, q = r & 15;
  return (k & ~31) + 8 * (q >> 2) + (q & 3) + 4 * (r >> 4);
}

// Column of element (row, col) of a table row of wd BF16 elements: the
// 16-byte chunks of a row are XOR-permuted by a key of the row so that the
// LDS reads of the MFMA operands spread over the 64 banks (4 bytes each).
// gfx950 serves a 16-byte read in four cycles of 16 lanes, lanes {0-3, 8-11,
// 20-23, 28-31} and {4-7, 12-15, 16-19, 24-27} of each half wave, and an
// 8-byte (transposing) read in two cycles of 32 lanes (bank-conflict
// counters of single-pattern kernels). An involution within the row; rows
// whose chunk count is not a power of two are left as they are.
//   * Rows of 256 bytes or more: key = row mod 16 with its bits 0..3 moved
//     to key bits 0, 2, 3, 1. A forward fragment read (rows 16 r + i, chunk
//     4 j + g) then puts each 16-lane cycle on 16 different chunks, and the
//     transposing reads of a W^T fragment (rows 4 g + k, one 8-byte half of
//     chunks 4 a + q) load every chunk of a 32-lane cycle exactly twice, the
//     least those reads allow (they all take the same half of their chunk).
//   * Rows of 128 bytes (two per 256 bytes of banks): key bit 1 = row bit 1,
//     key bit 2 = row bit 2 ^ row bit 3. Forward fragment reads put each
//     16-lane cycle on 16 different chunks of the two rows' 256 bytes, and
//     the transposing reads of W^T (rows 4 g + k, both halves of chunk pairs
//     2 t, 2 t + 1) put each 32-lane cycle on 64 different banks.
__host__ __device__ __forceinline__ int swz(int row, int col, int wd)
{
  const int nc = wd >> 3;
  if (nc & (nc - 1)) return col;
  const int r = row & 15;
  if (nc >= 16)
  {
    const int key = (r & 1) | ((r >> 2) & 2) | ((r << 1) & 4) | ((r << 1) & 8);
    return (((col >> 3) ^ key) << 3) | (col & 7);
  }
  if (nc == 8)
  {
    const int key = (r & 2) | ((((r >> 2) ^ (r >> 3)) & 1) << 2);
    return (((col >> 3) ^ key) << 3) | (col & 7);
  }
  const int per = 16 / nc;  // rows per 256 bytes
  return ((((col >> 3) ^ ((row / per) & (nc - 1)))) << 3) | (col & 7);
}

// Weight table layout (elements, BF16), identical in HBM and LDS; the columns
// of every row are placed by swz. The backward reads the transposed operands
// (W1^T, W0^T, Wout^T) from the same tables with transposing LDS reads.
struct TableLayout
{
  int H, FP, G;
  __host__ __device__ int oW0() const { return 0; }                   // [H][FP]
  __host__ __device__ int oW1() const { return H * FP; }              // [H][H]  cols kperm
  __host__ __device__ int oWo() const { return oW1() + H * H; }       // [16][H] cols kperm
  __host__ __device__ int total() const { return oWo() + 16 * H; }
};

// Addresses of the transposing reads that form the A fragment of a
// transposed table (the transposed M index is a table column, its K index a
// table row), for lane (g = lane / 16, i = lane % 16), half h of K chunk j:
// lane 4k + q reads row 32 j + 16 h + 4 g + k (the hidden unit at K position
// 32 j + 8 g + 4 h + k of the kperm order), logical columns 16 r + 4 q .. + 3.
__host__ __device__ __forceinline__ int tr_row(int j, int h, int g, int i) { return 32 * j + 16 * h + 4 * g + (i >> 2); }

struct FusedParams
{
  ntc_feature_desc fd;
  ntc_tiles tl;
  ntc_noise nz;
  const float* g0;
  const float* g1;
  const uint16_t* wtab;
  const uint8_t* pkg;
  int O;
  int F;
  float scale;
  int64_t B;
  int scatter;
  float* dg0;
  float* dg1;
  float* dw_part;
  double* sse_part;
  int dec_elems;
  float* yhat;         // forward: decoder outputs [O][B]
  const float* dyhat;  // backward: upstream gradient dL/dYhat [O][B]
};

// What a launch of the body does. TRAIN: the whole training step (the loss
// gradient from the targets). FORWARD: gather + forward only, writing the
// outputs (and the SSE partials when targets are given). BACKWARD: gather +
// forward again (recomputed, nothing is kept between the two launches), then
// the backward from an upstream gradient instead of the loss.
enum FusedMode : int
{
  kTrain = 0,
  kForward = 1,
  kBackward = 2,
};

// NT: 16-texel MFMA column tiles per wave; a workgroup pass covers 64*NT texels.
template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT, int NT>
struct Cfg
{
  static constexpr int PASS = 64 * NT;
  static constexpr int MT = H / 16;      // hidden M-tiles
  static constexpr int KF = FP / 32;     // K chunks over features
  static constexpr int KH = H / 32;      // K chunks over hidden units
  static constexpr int GT = G / 16;      // dX M-tiles (grid slots)
  static constexpr int RW = H / 4;       // dW rows per wave
  static constexpr int MTW = RW / 16;    // dW M-tiles per wave
  static constexpr int NTO = H / 64;     // dWout column tiles per wave
  static constexpr int G1B = G0MODE == NTC_G0_CORNERS ? 4 * C0 : C0;
  static constexpr int GSLOTS = G1B + C1;
  static constexpr int PE_FIRST = PEC == 3 ? 8 : 4;
  static constexpr int WT = H * FP + H * H + 16 * H;
  static constexpr int XS = FP + 8;  // X row stride (elements): spreads texel rows over LDS banks
  static constexpr int XR = PASS * XS;  // X rows [PASS][XS], kept from the gather to the rounds
  // Decoder-gradient staging of a 64-texel round, texel-major: A0, dZ1, dZ0
  // and A1 as [64][SH] BF16 rows (a texel's H values in the kperm position
  // order, one 16-byte store per fragment; SH = H + 16 puts the four rows
  // of a transposing read on distinct banks), dY as [64][16].
  static constexpr int SH = H + 16;
  static constexpr int STAGE = 4 * 64 * SH + 64 * 16;
  static constexpr int BH = PASS / 16;
  // Latent gradients of a blocked pass (16 x BH texels) come from the
  // per-texel dX when the pass footprint spans at most WMAX x VMAX cells of
  // the grid (texel:cell ratio >= 1 along both axes).
  static constexpr int WMAX = 18, VMAX = BH + 3;
  // Staged gather: the footprint cells of both grids, noise added, sit in
  // LDS as [v][u][CP] floats when the boxes span at most SW x SV cells;
  // channel quads (4 channels share one Philox draw) are prefetched, PF per
  // thread, while the previous pass computes its decoder gradients.
  static constexpr int Q0 = (C0 + 3) / 4, Q1 = (C1 + 3) / 4;
  static constexpr int CP0 = 4 * Q0, CP1 = 4 * Q1;
  static constexpr int PF = H >= 128 ? 2 : 3;
  // dX row stride = 16 mod 64 floats puts the 4 texel rows of one FP32 MFMA
  // B-operand read in distinct LDS banks
  static constexpr int GS_BANK = G + (((16 - G) % 64) + 64) % 64;
  static constexpr int lds_bytes(int sw, int sv, int gs, int stage)
  {
    const int u = 2 * sw * sv * (CP0 + CP1);
    return 2 * (WT + XR + (u > stage ? u : stage)) + 4 * PASS * gs + 4 * 192 + 4 * 128 + 16 * PASS;
  }
  static constexpr int LIM = 160 * 1024;
  static constexpr int SEL = lds_bytes(WMAX, BH + 3, GS_BANK, STAGE) <= LIM       ? 0
                             : lds_bytes(WMAX, BH + 3, G + 4, STAGE) <= LIM       ? 1
                             : lds_bytes(10, BH / 2 + 3, GS_BANK, STAGE) <= LIM   ? 2
                             : lds_bytes(10, BH / 2 + 3, G + 4, STAGE) <= LIM     ? 3
                                                                           : 4;
  static constexpr int SW = SEL <= 1 ? WMAX : (SEL <= 3 ? 10 : 0);
  static constexpr int SV = SEL <= 1 ? BH + 3 : (SEL <= 3 ? BH / 2 + 3 : 0);
  static constexpr int GS = (SEL == 0 || SEL == 2) ? GS_BANK : G + 4;  // dX row stride (floats)
  static constexpr int DXS = PASS * GS;
  static constexpr int SBOX0 = SW * SV * CP0;  // g1 box offset in the staged boxes (floats)
  static constexpr int UNION_U = 2 * SW * SV * (CP0 + CP1);
  static constexpr int UNION = UNION_U > STAGE ? UNION_U : STAGE;
  static constexpr int LDS = WT + XR + UNION;
  static_assert(FP % 32 == 0 && H % 64 == 0 && G % 16 == 0, "tile shapes");
  static_assert(GSLOTS <= G && G <= FP, "grid slots");
  static_assert(UNION * 2 >= 8 * 256, "final reduction reuses the union region");
  static_assert(16 + BH <= 128, "axis table threads");
  static_assert(((WT + XR) * 2) % 16 == 0, "staged boxes 16-byte aligned");
  static_assert(lds_bytes(SW, SV, GS, STAGE) <= LIM, "LDS budget");
  // box cell indices (below WMAX * VMAX + 16) divided by box widths (at most WMAX)
  static_assert((WMAX * VMAX + 16) * WMAX < 65536, "small_div range");
};

// Corner tables of one grid for a blocked pass: box-local corner columns
// and fractions of the pass's 16 texel columns and BH texel rows.
struct AxisTab
{
  float fx[16], fy[16];
  int8_t cx0[16], cx1[16], cy0[16], cy1[16];
};
static_assert(sizeof(AxisTab) == 192, "axis table size in the LDS budget");

// Where the texels of a workgroup pass lie. For tiles of side >= 16 a pass
// is a 16 x BH block of one tile (block-major order), so its latent footprint
// is compact; smaller tiles are walked linearly.
// log2 of a power of two
__device__ __forceinline__ int ilog2(int v) { return 31 - __builtin_clz((unsigned)v); }

// Division by a small runtime divisor d as a multiply and shift:
// q = (n * ceil(2^16 / d)) >> 16 equals n / d whenever n * d < 2^16
// (the rounding-up error of the reciprocal, below d / 2^16 per unit of n,
// then stays below the distance 1/d from n/d to the next integer).
__device__ __forceinline__ uint32_t small_recip(int d) { return (65535u + (uint32_t)d) / (uint32_t)d; }
__device__ __forceinline__ int small_div(int n, uint32_t recip) { return (int)(((uint32_t)n * recip) >> 16); }

struct PassMap
{
  bool blocked;
  int b, bx0, by0;
};

__device__ __forceinline__ PassMap pass_map(int64_t pass, int T, int pass_texels)
{
  PassMap pm;
  pm.blocked = T >= 16;
  pm.b = 0;
  pm.bx0 = 0;
  pm.by0 = 0;
  if (pm.blocked)
  {
    // T and the block height pass_texels / 16 are powers of two: a tile holds
    // T / 16 blocks per row and T / bh block rows
    const int lt = ilog2(T), lbh = ilog2(pass_texels) - 4;
    const int lrow = lt - 4, ltile = lrow + lt - lbh;
    pm.b = (int)(pass >> ltile);
    const int q = (int)(pass & (((int64_t)1 << ltile) - 1));
    pm.bx0 = (q & ((1 << lrow) - 1)) << 4;
    pm.by0 = (q >> lrow) << lbh;
  }
  return pm;
}

// Texel tl of the pass: mip position (x, y), unwrapped position (xu, yu) =
// origin + tile-local coordinate (>= M after a wrap), validity. orig holds
// the tile origins reduced mod M (unsigned: any origin value maps into
// [0, M)).
struct PassTexel
{
  int x, y, xu, yu;
  bool valid;
};

__device__ __forceinline__ PassTexel pass_texel(const PassMap& pm, int64_t pass, int tl, int pass_texels,
                                                const uint32_t* orig, int T, int M, int64_t B)
{
  PassTexel p;
  int b, lx, ly;
  if (pm.blocked)
  {
    b = pm.b;
    lx = pm.bx0 + (tl & 15);
    ly = pm.by0 + (tl >> 4);
    p.valid = true;
  }
  else
  {
    const int64_t t = pass * pass_texels + tl;
    p.valid = t < B;
    const int64_t tt = p.valid ? t : 0;
    const int lt = ilog2(T);  // T is a power of two
    b = (int)(tt >> (2 * lt));
    const int xy = (int)(tt & (((int64_t)1 << (2 * lt)) - 1));
    lx = xy & (T - 1);
    ly = xy >> lt;
  }
  p.xu = (int)(orig[2 * b] + (uint32_t)lx);
  p.yu = (int)(orig[2 * b + 1] + (uint32_t)ly);
  p.x = p.xu >= M ? p.xu - M : p.xu;
  p.y = p.yu >= M ? p.yu - M : p.yu;
  return p;
}

// Batch index of texel tl of the pass (tile-major, then rows of the tile, the
// order of the plain path's feature and target columns).
__device__ __forceinline__ int64_t pass_texel_index(const PassMap& pm, int64_t pass, int tl, int pass_texels, int T)
{
  if (!pm.blocked) return pass * pass_texels + tl;
  const int lx = pm.bx0 + (tl & 15), ly = pm.by0 + (tl >> 4);
  return ((int64_t)pm.b * T + ly) * T + lx;
}

// Cell box of one grid covering a blocked pass, in unwrapped cell coordinates.
struct Box
{
  int u0, v0, w, h;
  uint32_t wr;  // small_recip(w)
  bool fits;
};

__device__ __forceinline__ Box make_box(int xu0, int xu1, int yu0, int yu1, int M, int D, int wmax, int vmax)
{
  Box bx;
  int a, b;
  float f;
  axis_corners(xu0, M, D, 0, bx.u0, a, f);
  axis_corners(xu1, M, D, 0, a, b, f);
  bx.w = b - bx.u0 + 1;
  axis_corners(yu0, M, D, 0, bx.v0, a, f);
  axis_corners(yu1, M, D, 0, a, b, f);
  bx.h = b - bx.v0 + 1;
  bx.fits = bx.w <= wmax && bx.h <= vmax;
  bx.wr = small_recip(bx.w);
  return bx;
}

// Fills the corner tables of one grid; t in [0, 128) is the thread's slot.
template <int BH>
__device__ __forceinline__ void fill_axis_tab(AxisTab& tb, const Box& bx, int xu0, int yu0, int M, int D, int t)
{
  int c0, c1;
  float f;
  if (t < 16)
  {
    axis_corners(xu0 + t, M, D, bx.u0, c0, c1, f);
    tb.cx0[t] = (int8_t)c0;
    tb.cx1[t] = (int8_t)c1;
    tb.fx[t] = f;
  }
  else if (t < 16 + BH)
  {
    const int r = t - 16;
    axis_corners(yu0 + r, M, D, bx.v0, c0, c1, f);
    tb.cy0[r] = (int8_t)c0;
    tb.cy1[r] = (int8_t)c1;
    tb.fy[r] = f;
  }
}

__device__ __forceinline__ f32x4 mfma_f32(float a, float b, const f32x4& c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Gradient of every (cell, channel) of a grid's box as one FP32 MFMA product
// per 16 cells x 16 channels: S[cell][ch] = sum over the pass's texels t of
// Wt[cell][t] * dX[t][slot(ch)], where Wt holds the gather's corner weights
// ((1-fx or fx) * (1-fy or fy) for bilinear grids; 1 for the unweighted
// corner layout of g0, one product per corner c reading slot c * C + ch).
// Each sum goes to the grid's gradient with one global atomic.
// MFMA 16x16x4 operands: lane l gives A[l%16][l/16] and B[l/16][l%16] and
// receives D[4(l/16)+i][l%16]; K step kk covers texels 4kk..4kk+3, so lane l
// meets texel column 4(kk%4) + l/16 of texel row kk/4.
template <int C, bool CORNERS, int GS, int PASS>
__device__ __forceinline__ void box_gradients(const AxisTab& tb, const Box& bx, const float* dxs, int slot0,
                                              float* dg, int D, int wave, int lane)
{
  constexpr int BH = PASS / 16;
  constexpr int KS = PASS / 4;
  constexpr int NTC = (C + 15) / 16;
  constexpr int NC = CORNERS ? 4 : 1;
  const int cells = bx.w * bx.h;
  const int tiles = ((cells + 15) / 16) * NTC;
  if (wave >= tiles) return;
  const int q = lane >> 4, n = lane & 15;
  int ax0[4], ax1[4], ay0[BH], ay1[BH];
  float fx[4], fy[BH];
#pragma unroll
  for (int j = 0; j < 4; ++j)
  {
    ax0[j] = tb.cx0[4 * j + q];
    ax1[j] = tb.cx1[4 * j + q];
    fx[j] = tb.fx[4 * j + q];
  }
#pragma unroll
  for (int y = 0; y < BH; ++y)
  {
    ay0[y] = tb.cy0[y];
    ay1[y] = tb.cy1[y];
    fy[y] = tb.fy[y];
  }
  for (int tile = wave; tile < tiles; tile += 4)
  {
    const int mt = tile / NTC, ct = tile - mt * NTC;
    const int m = 16 * mt + n;
    const int v = small_div(m, bx.wr), u = m - v * bx.w;  // v >= bx.h past the last cell: no corner matches
    const int ch = 16 * ct + n;
    // two accumulators (even and odd K steps) so that consecutive MFMAs do
    // not wait for each other; summed at the end
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc;
#pragma unroll
    for (int c = 0; c < NC; ++c)
    {
      // x and y factors of this lane's cell; one of the two terms is zero
      // unless both corners clamp onto the same border cell
      float wx[4], wy[BH];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        wx[j] = CORNERS ? (((c & 1) ? ax1[j] : ax0[j]) == u ? 1.0f : 0.0f)
                        : (ax0[j] == u ? 1.0f - fx[j] : 0.0f) + (ax1[j] == u ? fx[j] : 0.0f);
#pragma unroll
      for (int y = 0; y < BH; ++y)
        wy[y] = CORNERS ? (((c >> 1) ? ay1[y] : ay0[y]) == v ? 1.0f : 0.0f)
                        : (ay0[y] == v ? 1.0f - fy[y] : 0.0f) + (ay1[y] == v ? fy[y] : 0.0f);
#pragma unroll
      for (int kk = 0; kk < KS; ++kk)
      {
        const float d = ch < C ? dxs[(4 * kk + q) * GS + slot0 + c * C + ch] : 0.0f;
        if (kk & 1)
          acc1 = mfma_f32(wx[kk & 3] * wy[kk >> 2], d, acc1);
        else
          acc = mfma_f32(wx[kk & 3] * wy[kk >> 2], d, acc);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] += acc1[i];
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      const int cell = 16 * mt + 4 * q + i;
      if (cell < cells && ch < C)
      {
        const int rv = small_div(cell, bx.wr);
        const int cu = bx.u0 + cell - rv * bx.w, cv = bx.v0 + rv;
        const int cx = cu >= D ? cu - D : cu, cy = cv >= D ? cv - D : cv;
        atomicAdd(&dg[((int64_t)cy * D + cx) * C + ch], acc[i]);
      }
    }
  }
}

// Geometry of one workgroup pass: where its texels lie, the cell boxes of
// both grids and whether its gather is staged through LDS.
struct PassGeo
{
  PassMap pm;
  int xu0, yu0;  // unwrapped mip position of the block's first texel
  Box b0, b1;
  bool staged;
};

template <class CF>
__device__ __forceinline__ PassGeo pass_geo(int64_t pass, const uint32_t* orig, int T, int M, int D0, int D1)
{
  PassGeo pg;
  pg.pm = pass_map(pass, T, CF::PASS);
  pg.xu0 = 0;
  pg.yu0 = 0;
  pg.b0 = Box{0, 0, 0, 0, 65536u, false};
  pg.b1 = Box{0, 0, 0, 0, 65536u, false};
  pg.staged = false;
  if (pg.pm.blocked)
  {
    pg.xu0 = (int)orig[2 * pg.pm.b] + pg.pm.bx0;
    pg.yu0 = (int)orig[2 * pg.pm.b + 1] + pg.pm.by0;
    pg.b0 = make_box(pg.xu0, pg.xu0 + 15, pg.yu0, pg.yu0 + CF::BH - 1, M, D0, CF::WMAX, CF::VMAX);
    pg.b1 = make_box(pg.xu0, pg.xu0 + 15, pg.yu0, pg.yu0 + CF::BH - 1, M, D1, CF::WMAX, CF::VMAX);
    pg.staged = pg.b0.w <= CF::SW && pg.b0.h <= CF::SV && pg.b1.w <= CF::SW && pg.b1.h <= CF::SV &&
                pg.b0.w * pg.b0.h * CF::Q0 + pg.b1.w * pg.b1.h * CF::Q1 <= CF::PF * kThreads;
  }
  return pg;
}

// Channel quads of a staged pass's footprint held in registers between the
// global loads and their noisy copy into LDS.
template <int PF>
struct Prefetch
{
  float v[PF][4];
  int dst[PF];   // float offset in the staged boxes; -1: nothing
  int cell[PF];  // grid cell (noise counter)
  int gq[PF];    // grid | quad << 1
};

// Issues the loads of a pass's staged footprint quads (PF per thread) into
// registers. Every load is unconditional, from a valid address (idle slots
// read the first quad), so the issue point stays where it is written.
template <class CF, int C0, int C1>
__device__ __forceinline__ void prefetch_issue(Prefetch<CF::PF>& pf, const PassGeo& pg, const FusedParams& P, int tid)
{
  const int n0 = pg.b0.w * pg.b0.h * CF::Q0;
  const int n = pg.staged ? n0 + pg.b1.w * pg.b1.h * CF::Q1 : 0;
#pragma unroll
  for (int k = 0; k < CF::PF; ++k)
  {
    const int e0 = tid + k * kThreads;
    const bool live = e0 < n;
    const bool g1 = live && e0 >= n0;
    const int ee = !live ? 0 : (g1 ? e0 - n0 : e0);
    const int Q = g1 ? CF::Q1 : CF::Q0, C = g1 ? C1 : C0, CP = g1 ? CF::CP1 : CF::CP0;
    const int bw = g1 ? pg.b1.w : pg.b0.w, bu0 = g1 ? pg.b1.u0 : pg.b0.u0, bv0 = g1 ? pg.b1.v0 : pg.b0.v0;
    const uint32_t bwr = g1 ? pg.b1.wr : pg.b0.wr;
    const int D = g1 ? P.fd.g1_dim : P.fd.g0_dim;
    const int q = ee % Q, cb = ee / Q;
    const int v = small_div(cb, bwr), u = cb - v * bw;
    const int cu = bu0 + u, cv = bv0 + v;
    const int cx = cu >= D ? cu - D : cu, cy = cv >= D ? cv - D : cv;
    const int cell = cy * D + cx;
    const float* src = (g1 ? P.g1 : P.g0) + (int64_t)cell * C + 4 * q;
#pragma unroll
    for (int i = 0; i < 4; ++i) pf.v[k][i] = src[(4 * q + i < C) ? i : 0];
    pf.dst[k] = live ? (g1 ? CF::SBOX0 : 0) + cb * CP + 4 * q : -1;
    pf.cell[k] = cell;
    pf.gq[k] = (int)g1 | (q << 1);
  }
}

// Quantisation noise added per channel quad as in the per-texel gather
// (one Philox draw per cell and quad), then stored into the staged boxes.
template <class CF, int C0, int C1>
__device__ __forceinline__ void prefetch_commit(const Prefetch<CF::PF>& pf, const FusedParams& P, float* sbox)
{
#pragma unroll
  for (int k = 0; k < CF::PF; ++k)
  {
    if (pf.dst[k] < 0) continue;
    const int grid = pf.gq[k] & 1, q = pf.gq[k] >> 1;
    const int C = grid ? C1 : C0;
    const float delta = grid ? P.nz.delta1 : P.nz.delta0;
    uint4 w = make_uint4(0, 0, 0, 0);
    if (P.nz.enabled) w = noise_words(P.nz.seed, P.nz.step, grid, P.nz.level, (uint32_t)pf.cell[k], q);
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      float v = pf.v[k][i];
      if (P.nz.enabled) v = v + noise_value(pick(w, i), delta);
      if (4 * q + i < C) sbox[pf.dst[k] + i] = v;
    }
  }
}

// Issues the loads of the target bytes of this lane's output rows
// (4(l/16)+i) for its texels of a pass; unconditional, like prefetch_issue.
template <class CF, int NT, class PP>
__device__ __forceinline__ void targets_issue(uint32_t (&tg)[NT][4], const PassMap& pm, int64_t pass,
                                              const PP& P, const uint32_t* orig, int wave, int n16, int g)
{
  const int M = P.tl.mip_dim;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
  {
    const PassTexel pt = pass_texel(pm, pass, wave * 16 * NT + nt * 16 + n16, CF::PASS, orig, P.tl.tile_dim, M, P.B);
    const int64_t base = pt.valid ? ((int64_t)pt.y * M + pt.x) * P.O : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      const int o = 4 * g + i;
      tg[nt][i] = (uint32_t)P.pkg[base + (o < P.O ? o : 0)];
    }
  }
}

// Target bytes into LDS, one word per (texel, lane group): [PASS][4] words.
template <int NT>
__device__ __forceinline__ void targets_commit(const uint32_t (&tg)[NT][4], uint32_t* tgt_s, int wave, int n16, int g)
{
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
    tgt_s[(wave * 16 * NT + nt * 16 + n16) * 4 + g] = tg[nt][0] | (tg[nt][1] << 8) | (tg[nt][2] << 16) | (tg[nt][3] << 24);
}

// Feature slots after the grids: positional encoding, lod, zero padding.
template <class CF, int PEC, int LODS, class Put>
__device__ __forceinline__ void gather_rest(const FusedParams& P, int px, int py, Put&& put)
{
  constexpr int FP = CF::KF * 32;
  int slot = CF::GSLOTS;
#pragma unroll
  for (int pi = 0; pi < PEC; ++pi)
  {
    const int Pd = CF::PE_FIRST >> pi;
#pragma unroll
    for (int axis = 0; axis < 2; ++axis)
#pragma unroll
      for (int h = 0; h < 2; ++h) put(slot++, tri_wave<float>(axis == 0 ? px : py, Pd, h));
  }
  if (LODS) put(slot++, P.tl.norm_lod);
  for (; slot < FP; ++slot) put(slot, 0.0f);
}

// Features of a staged pass from the LDS boxes: one item per (channel quad
// of a grid, texel) plus one per texel for the remaining slots. Corner
// weights and the fma order are those of the per-texel gather.
template <class CF, int C0, int C1, int G0MODE, int PEC, int LODS>
__device__ __forceinline__ void gather_staged(const FusedParams& P, const PassGeo& pg, const AxisTab* tab,
                                              const float* sbox, int tid, uint16_t* xnf)
{
  constexpr int Q0 = CF::Q0, Q1 = CF::Q1, QI = Q0 + Q1 + 1;
  const int M = P.tl.mip_dim;
  for (int item = tid; item < QI * CF::PASS; item += kThreads)
  {
    const int qi = item / CF::PASS, t = item % CF::PASS;
    const int x = t & 15, y = t >> 4;
    auto put = [&](int slot, float v) { xnf[t * CF::XS + slot] = bf16_dev(v); };
    // four consecutive slots from slot0 (a multiple of 4): one 8-byte write
    // into the texel's X row (rows of XS elements, XS a multiple of 4)
    auto put4 = [&](int slot0, float a, float b, float c, float d) {
      *reinterpret_cast<uint2*>(&xnf[t * CF::XS + slot0]) = make_uint2(pk_bf16(a, b), pk_bf16(c, d));
    };
    if (qi < Q0 + Q1)
    {
      const bool g1 = qi >= Q0;
      const int q = g1 ? qi - Q0 : qi;
      const AxisTab& tb = tab[g1 ? 1 : 0];
      const int bw = g1 ? pg.b1.w : pg.b0.w;
      const int CP = g1 ? CF::CP1 : CF::CP0, C = g1 ? C1 : C0;
      const float* box = sbox + (g1 ? CF::SBOX0 : 0) + 4 * q;
      const int ax0 = tb.cx0[x], ax1 = tb.cx1[x], ay0 = tb.cy0[y], ay1 = tb.cy1[y];
      const float fx = tb.fx[x], fy = tb.fy[y];
      const float w[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
      const int off[4] = {(ay0 * bw + ax0) * CP, (ay0 * bw + ax1) * CP, (ay1 * bw + ax0) * CP,
                          (ay1 * bw + ax1) * CP};
      float4 val[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) val[c] = *reinterpret_cast<const float4*>(box + off[c]);
      if (!g1 && G0MODE == NTC_G0_CORNERS)
      {
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
          if (C0 % 4 == 0)
            put4(c * C0 + 4 * q, val[c].x, val[c].y, val[c].z, val[c].w);
          else
          {
            if (4 * q + 0 < C0) put(c * C0 + 4 * q + 0, val[c].x);
            if (4 * q + 1 < C0) put(c * C0 + 4 * q + 1, val[c].y);
            if (4 * q + 2 < C0) put(c * C0 + 4 * q + 2, val[c].z);
            if (4 * q + 3 < C0) put(c * C0 + 4 * q + 3, val[c].w);
          }
        }
      }
      else
      {
        const int base = (g1 ? CF::G1B : 0) + 4 * q;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
          acc[0] = fmaf(w[c], val[c].x, acc[0]);
          acc[1] = fmaf(w[c], val[c].y, acc[1]);
          acc[2] = fmaf(w[c], val[c].z, acc[2]);
          acc[3] = fmaf(w[c], val[c].w, acc[3]);
        }
        if (C0 % 4 == 0 && C1 % 4 == 0 && CF::G1B % 4 == 0)
          put4(base, acc[0], acc[1], acc[2], acc[3]);
        else
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (4 * q + i < C) put(base + i, acc[i]);
      }
    }
    else
    {
      const int xu = pg.xu0 + x, yu = pg.yu0 + y;
      if constexpr (C0 % 4 == 0 && C1 % 4 == 0 && CF::G1B % 4 == 0)
      {
        // the remaining slots (positional encoding, lod, zero padding) in
        // groups of four
        constexpr int NR = CF::KF * 32 - CF::GSLOTS;
        float v[NR];
        gather_rest<CF, PEC, LODS>(P, xu >= M ? xu - M : xu, yu >= M ? yu - M : yu,
                                   [&](int slot, float f) { v[slot - CF::GSLOTS] = f; });
#pragma unroll
        for (int r = 0; r < NR; r += 4) put4(CF::GSLOTS + r, v[r], v[r + 1], v[r + 2], v[r + 3]);
      }
      else
        gather_rest<CF, PEC, LODS>(P, xu >= M ? xu - M : xu, yu >= M ? yu - M : yu, put);
    }
  }
}

// Gather one texel's feature slots into the LDS copies of X directly from the
// grids (passes whose footprint is not staged): part 0 the g0 slots, part 1
// g1 and the remaining slots.
template <class CF, int C0, int C1, int G0MODE, int PEC, int LODS>
__device__ __forceinline__ void gather_part(const FusedParams& P, int part, int tl, bool valid, int px, int py,
                                            uint16_t* xnf)
{
  constexpr int FP = CF::KF * 32;
  auto put = [&](int slot, float v) { xnf[tl * CF::XS + slot] = bf16_dev(v); };
  const int M = P.tl.mip_dim;
  if (part == 0)
  {
    if (!valid)
    {
      for (int s = 0; s < CF::G1B; ++s) put(s, 0.0f);
      return;
    }
    const Footprint<float> f0 = footprint<float>(px, py, M, P.fd.g0_dim);
    if (G0MODE == NTC_G0_CORNERS)
    {
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
#pragma unroll
        for (int q = 0; q < (C0 + 3) / 4; ++q)
        {
          const uint4 w = P.nz.enabled ? noise_words(P.nz.seed, P.nz.step, 0, P.nz.level, (uint32_t)f0.cell[c], q)
                                       : make_uint4(0, 0, 0, 0);
#pragma unroll
          for (int e = 0; e < 4; ++e)
          {
            const int ch = 4 * q + e;
            if (ch < C0)
            {
              float v = P.g0[(int64_t)f0.cell[c] * C0 + ch];
              if (P.nz.enabled) v = v + noise_value(pick(w, e), P.nz.delta0);
              put(c * C0 + ch, v);
            }
          }
        }
      }
    }
    else
    {
#pragma unroll
      for (int q = 0; q < (C0 + 3) / 4; ++q)
      {
        uint4 w[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
          w[c] = P.nz.enabled ? noise_words(P.nz.seed, P.nz.step, 0, P.nz.level, (uint32_t)f0.cell[c], q)
                              : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
          const int ch = 4 * q + e;
          if (ch < C0)
          {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
              float v = P.g0[(int64_t)f0.cell[c] * C0 + ch];
              if (P.nz.enabled) v = v + noise_value(pick(w[c], e), P.nz.delta0);
              acc = fmaf(f0.w[c], v, acc);
            }
            put(ch, acc);
          }
        }
      }
    }
    return;
  }
  // part 1: g1, positional encoding, lod, zero padding
  if (!valid)
  {
    for (int s = CF::G1B; s < FP; ++s) put(s, 0.0f);
    return;
  }
  const Footprint<float> f1 = footprint<float>(px, py, M, P.fd.g1_dim);
#pragma unroll
  for (int q = 0; q < (C1 + 3) / 4; ++q)
  {
    uint4 w[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      w[c] = P.nz.enabled ? noise_words(P.nz.seed, P.nz.step, 1, P.nz.level, (uint32_t)f1.cell[c], q)
                          : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
      const int ch = 4 * q + e;
      if (ch < C1)
      {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
          float v = P.g1[(int64_t)f1.cell[c] * C1 + ch];
          if (P.nz.enabled) v = v + noise_value(pick(w[c], e), P.nz.delta1);
          acc = fmaf(f1.w[c], v, acc);
        }
        put(CF::G1B + ch, acc);
      }
    }
  }
  gather_rest<CF, PEC, LODS>(P, px, py, put);
}

// One workgroup's share of a step: passes vblk, vblk + vgrid, ... of the
// batch described by P; its decoder-gradient and SSE partials go to slot vblk.
template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT, int NT, int MODE = kTrain>
__device__ __forceinline__ void fused_body(const FusedParams& P, const int vblk, const int vgrid)
{
  using CF = Cfg<H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT, NT>;
  constexpr int kNT = NT, kPass = CF::PASS;
  constexpr int MT = CF::MT, KF = CF::KF, KH = CF::KH, GT = CF::GT, MTW = CF::MTW, NTO = CF::NTO;
  constexpr int kRing = 8;  // fragments read ahead of their MFMA
  constexpr TableLayout TLY = {H, FP, G};
  __shared__ __attribute__((aligned(16))) uint16_t smem[CF::LDS];
  uint16_t* wt = smem;
  uint16_t* xnf = smem + CF::WT;                       // X rows [PASS][XS]
  uint16_t* uni = xnf + CF::XR;
  float* sbox = reinterpret_cast<float*>(uni);         // staged boxes during the gather
  // decoder-gradient staging of a round: texel rows
  uint16_t* stA = uni;                       // [64][SH] A0
  uint16_t* stB = uni + 64 * CF::SH;         // [64][SH] dZ1
  uint16_t* stC = uni + 2 * 64 * CF::SH;     // [64][SH] dZ0
  uint16_t* stD = uni + 3 * 64 * CF::SH;     // [64][SH] A1
  uint16_t* stY = uni + 4 * 64 * CF::SH;     // [64][16] dY
  __shared__ __attribute__((aligned(16))) float dxs[CF::DXS];  // dX of the grid slots [texel][GS]
  __shared__ AxisTab atab[4];                                  // corner tables of g0, g1; two passes
  __shared__ uint32_t orig_s[128];                             // tile origins mod M
  __shared__ uint32_t tgt_s[kPass * 4];                        // target bytes [texel][lane group]

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int n16 = lane & 15, g = lane >> 4;

  // weights -> LDS (16-byte copies)
  {
    const uint4* src = reinterpret_cast<const uint4*>(P.wtab);
    uint4* dst = reinterpret_cast<uint4*>(wt);
    for (int i = tid; i < CF::WT / 8; i += kThreads) dst[i] = src[i];
  }

  f32x4 dw1[MTW][MT], dw0[MTW][FP / 16], dwo[NTO];
#pragma unroll
  for (int a = 0; a < MTW; ++a)
  {
#pragma unroll
    for (int c = 0; c < MT; ++c) dw1[a][c] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < FP / 16; ++c) dw0[a][c] = f32x4{0, 0, 0, 0};
  }
#pragma unroll
  for (int c = 0; c < NTO; ++c) dwo[c] = f32x4{0, 0, 0, 0};
  double sse = 0.0;

  const int64_t npass = (P.B + kPass - 1) / kPass;
  const bool use_targets = MODE == kTrain || (MODE == kForward && P.pkg != nullptr);
  const int T = P.tl.tile_dim, M = P.tl.mip_dim;
  const int D0 = P.fd.g0_dim, D1 = P.fd.g1_dim;
  if (tid < 2 * P.tl.n_tiles) orig_s[tid] = (uint32_t)P.tl.origins[tid] % (uint32_t)M;
  __syncthreads();

  // Corner tables, staged footprint and target bytes of a pass are written
  // into LDS by the previous pass, after its decoder-gradient rounds (their
  // loads are issued before the rounds); the first pass's here.
  auto fill_tables = [&](const PassGeo& pg, AxisTab* tab) {
    if (pg.pm.blocked)
    {
      if (tid < 128)
      {
        if (pg.b0.fits) fill_axis_tab<CF::BH>(tab[0], pg.b0, pg.xu0, pg.yu0, M, D0, tid);
      }
      else if (pg.b1.fits)
        fill_axis_tab<CF::BH>(tab[1], pg.b1, pg.xu0, pg.yu0, M, D1, tid - 128);
    }
  };
  PassGeo cur = pass_geo<CF>((int64_t)vblk, orig_s, T, M, D0, D1);
  int tb = 0;  // table pair of the current pass
  {
    Prefetch<CF::PF> pf;
    uint32_t tg[kNT][4];
    prefetch_issue<CF, C0, C1>(pf, cur, P, tid);
    if (use_targets) targets_issue<CF>(tg, cur.pm, (int64_t)vblk, P, orig_s, wave, n16, g);
    fill_tables(cur, atab);
    prefetch_commit<CF, C0, C1>(pf, P, sbox);
    if (use_targets) targets_commit(tg, tgt_s, wave, n16, g);
  }

  for (int64_t pass = vblk; pass < npass; pass += vgrid)
  {
    __syncthreads();  // this pass's tables, staged boxes and targets are in LDS
    const PassMap pm = cur.pm;
    const Box bx0 = cur.b0, bx1 = cur.b1;
    const AxisTab* tab = atab + 2 * tb;
    // ---- P0: gather ----
    if (cur.staged)
    {
      gather_staged<CF, C0, C1, G0MODE, PEC, LODS>(P, cur, tab, sbox, tid, xnf);
    }
    else
    {
      // threads [0,128) gather g0 slots, [128,256) the rest
      const int part = tid >> 7;
      for (int tl = tid & 127; tl < kPass; tl += 128)
      {
        const PassTexel pt = pass_texel(pm, pass, tl, kPass, orig_s, T, M, P.B);
        gather_part<CF, C0, C1, G0MODE, PEC, LODS>(P, part, tl, pt.valid, pt.x, pt.y, xnf);
      }
    }
    __syncthreads();

    // ---- P1: per-wave forward + backward for 32 texels ----
    int tloc[kNT];
    bool tvalid[kNT];
    int tx[kNT], ty[kNT];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt)
    {
      tloc[nt] = wave * 16 * kNT + nt * 16 + n16;
      const PassTexel pt = pass_texel(pm, pass, tloc[nt], kPass, orig_s, T, M, P.B);
      tvalid[nt] = pt.valid;
      tx[nt] = pt.x;
      ty[nt] = pt.y;
    }
    // layer 0
    f32x4 z0[MT][kNT];
    {
      bf16x8 xb[KF][kNT];
#pragma unroll
      for (int j = 0; j < KF; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt) xb[j][nt] = frag_lds(&xnf[tloc[nt] * CF::XS + 32 * j + 8 * g]);
      lds_ring<MT * KF, kRing>(
          [&](int i) {
            const int row = 16 * (i % MT) + n16;
            return &wt[TLY.oW0() + row * FP + swz(row, 32 * (i / MT) + 8 * g, FP)];
          },
          [&](int i, const bf16x8 a) {
            const int r = i % MT, j = i / MT;
            if (j == 0)
#pragma unroll
              for (int nt = 0; nt < kNT; ++nt) z0[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) z0[r][nt] = mfma(a, xb[j][nt], z0[r][nt]);
          });
    }
    bf16x8 a0f[KH][kNT];
#pragma unroll
    for (int j = 0; j < KH; ++j)
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) a0f[j][nt] = act_frag<ACT>(z0[2 * j][nt], z0[2 * j + 1][nt]);
    // layer 1
    f32x4 z1[MT][kNT];
    lds_ring<MT * KH, kRing>(
        [&](int i) {
          const int row = 16 * (i % MT) + n16;
          return &wt[TLY.oW1() + row * H + swz(row, 32 * (i / MT) + 8 * g, H)];
        },
        [&](int i, const bf16x8 a) {
          const int r = i % MT, j = i / MT;
          if (j == 0)
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) z1[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
          for (int nt = 0; nt < kNT; ++nt) z1[r][nt] = mfma(a, a0f[j][nt], z1[r][nt]);
        });
    bf16x8 a1f[KH][kNT];
#pragma unroll
    for (int j = 0; j < KH; ++j)
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) a1f[j][nt] = act_frag<ACT>(z1[2 * j][nt], z1[2 * j + 1][nt]);
    // output layer + loss
    float dy[kNT][4];
    bf16x8 dyf[kNT];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt)
    {
      f32x4 yh = {0, 0, 0, 0};
      lds_ring<KH, kRing>([&](int j) { return &wt[TLY.oWo() + n16 * H + swz(n16, 32 * j + 8 * g, H)]; },
                          [&](int j, const bf16x8 a) { yh = mfma(a, a1f[j][nt], yh); });
      float v[8];
#pragma unroll
      for (int i = 0; i < 4; ++i)
      {
        const int o = 4 * g + i;
        float d = 0.0f;
        if (tvalid[nt] && o < P.O)
        {
          if constexpr (MODE == kBackward)
            d = P.dyhat[(int64_t)o * P.B + pass_texel_index(pm, pass, tloc[nt], kPass, T)];
          else
          {
            if constexpr (MODE == kForward) P.yhat[(int64_t)o * P.B + pass_texel_index(pm, pass, tloc[nt], kPass, T)] = yh[i];
            if (use_targets)
            {
              const float yt = byte_over_255((tgt_s[tloc[nt] * 4 + g] >> (8 * i)) & 0xFFu);
              const float r = yt - yh[i];
              sse += (double)r * (double)r;
              d = P.scale * r;
            }
          }
        }
        dy[nt][i] = d;
        v[i] = d;
        v[4 + i] = 0.0f;
      }
      dyf[nt] = frag_pack(v);
    }
    // backward through layer 2 and the activation of layer 1
    bf16x8 dz1f[KH][kNT];
    if constexpr (MODE != kForward)
    {
      f32x4 dz1[MT][kNT];
      lds_ring_ld<MT, kRing>(
          [&](int r) {
            const int o = 4 * g + (n16 >> 2), c = kperm_col(16 * r + 4 * (n16 & 3));
            return frag_tr_half(&wt[TLY.oWo() + o * H + swz(o, c, H)]);
          },
          [&](int r, const bf16x8 a) {
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) dz1[r][nt] = mfma(a, dyf[nt], f32x4{0, 0, 0, 0});
          });
#pragma unroll
      for (int j = 0; j < KH; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
          dz1f[j][nt] = act_bwd_frag<ACT>(dz1[2 * j][nt], dz1[2 * j + 1][nt], a1f[j][nt], z1[2 * j][nt],
                                          z1[2 * j + 1][nt]);
    }
    // backward through layer 1 and the activation of layer 0
    bf16x8 dz0f[KH][kNT];
    if constexpr (MODE != kForward)
    {
      f32x4 dz0[MT][kNT];
      lds_ring_ld<MT * KH, kRing>(
          [&](int i) {
            const int j = i / MT, c = kperm_col(16 * (i % MT) + 4 * (n16 & 3));
            const int r0 = tr_row(j, 0, g, n16), r1 = tr_row(j, 1, g, n16);
            return frag_tr(&wt[TLY.oW1() + r0 * H + swz(r0, c, H)], &wt[TLY.oW1() + r1 * H + swz(r1, c, H)]);
          },
          [&](int i, const bf16x8 a) {
            const int r = i % MT, j = i / MT;
            if (j == 0)
#pragma unroll
              for (int nt = 0; nt < kNT; ++nt) dz0[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) dz0[r][nt] = mfma(a, dz1f[j][nt], dz0[r][nt]);
          });
#pragma unroll
      for (int j = 0; j < KH; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
          dz0f[j][nt] = act_bwd_frag<ACT>(dz0[2 * j][nt], dz0[2 * j + 1][nt], a0f[j][nt], z0[2 * j][nt],
                                          z0[2 * j + 1][nt]);
    }
    // dX for the grid slots: kept per texel in LDS for the box gradients of
    // grids whose pass footprint fits a box, else scattered per corner with
    // global atomics
    if (MODE != kForward && P.scatter)
    {
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
      {
        // Every lane runs the MFMAs (lane l supplies row l%16 of the A
        // operand for the whole wave); only the writes are per texel.
        const bool tv = tvalid[nt];
        const Footprint<float> f0 = footprint<float>(tx[nt], ty[nt], M, D0);
        const Footprint<float> f1 = footprint<float>(tx[nt], ty[nt], M, D1);
        float* drow = dxs + tloc[nt] * CF::GS;
        f32x4 dx = {0, 0, 0, 0};
        lds_ring_ld<GT * KH, kRing>(
            [&](int q) {
              const int j = q % KH, c = 16 * (q / KH) + 4 * (n16 & 3);
              const int r0 = tr_row(j, 0, g, n16), r1 = tr_row(j, 1, g, n16);
              return frag_tr(&wt[TLY.oW0() + r0 * FP + swz(r0, c, FP)], &wt[TLY.oW0() + r1 * FP + swz(r1, c, FP)]);
            },
            [&](int q, const bf16x8 a) {
          const int r = q / KH, j = q % KH;
          if (j == 0) dx = f32x4{0, 0, 0, 0};
          dx = mfma(a, dz0f[j][nt], dx);
          if (j != KH - 1 || !tv) return;
          // slots 16 r + 4 g .. + 3 lie in one grid's range when its slot
          // counts are multiples of 4: one 16-byte write of the texel's dX row
          if constexpr (CF::G1B % 4 == 0 && CF::GSLOTS % 4 == 0)
          {
            const int s0 = 16 * r + 4 * g;
            if ((s0 < CF::G1B && bx0.fits) || (s0 >= CF::G1B && s0 < CF::GSLOTS && bx1.fits))
            {
              *reinterpret_cast<f32x4*>(drow + s0) = dx;
              return;
            }
          }
#pragma unroll
          for (int i = 0; i < 4; ++i)
          {
            const int slot = 16 * r + 4 * g + i;
            if (slot < CF::G1B)
            {
              if (bx0.fits)
                drow[slot] = dx[i];
              else if (G0MODE == NTC_G0_CORNERS)
              {
                const int c = slot / C0, ch = slot - (slot / C0) * C0;
                // corner c is lane-dependent: select it instead of indexing
                const int cell = c == 0 ? f0.cell[0] : (c == 1 ? f0.cell[1] : (c == 2 ? f0.cell[2] : f0.cell[3]));
                atomicAdd(&P.dg0[(int64_t)cell * C0 + ch], dx[i]);
              }
              else
              {
#pragma unroll
                for (int c = 0; c < 4; ++c) atomicAdd(&P.dg0[(int64_t)f0.cell[c] * C0 + slot], f0.w[c] * dx[i]);
              }
            }
            else if (slot < CF::GSLOTS)
            {
              if (bx1.fits)
                drow[slot] = dx[i];
              else
              {
                const int ch = slot - CF::G1B;
#pragma unroll
                for (int c = 0; c < 4; ++c) atomicAdd(&P.dg1[(int64_t)f1.cell[c] * C1 + ch], f1.w[c] * dx[i]);
              }
            }
          }
            });
      }
    }

    __syncthreads();  // everyone done reading xnf; dX rows complete
    // next pass's footprint and targets: the loads land during the rounds below
    const int64_t next = pass + vgrid;
    const int64_t pn = next < npass ? next : pass;
    PassGeo nxt = pass_geo<CF>(pn, orig_s, T, M, D0, D1);
    nxt.staged = nxt.staged && next < npass;
    Prefetch<CF::PF> pf;
    uint32_t tg[kNT][4];
    prefetch_issue<CF, C0, C1>(pf, nxt, P, tid);
    if (use_targets) targets_issue<CF>(tg, nxt.pm, pn, P, orig_s, wave, n16, g);
#pragma unroll
    for (int rnd = 0; rnd < (MODE == kForward ? 0 : kNT); ++rnd)
    {
      // this wave's tile rnd: texel rows wave * 16 + n16 of the round, one
      // 16-byte store per fragment (positions 32 j + 8 g .. + 7), dY as 4
      // outputs 4 g .. 4 g + 3
      {
        const int row = wave * 16 + n16;
#pragma unroll
        for (int j = 0; j < KH; ++j)
        {
          const int at = row * CF::SH + 32 * j + 8 * g;
          *reinterpret_cast<bf16x8*>(&stA[at]) = a0f[j][rnd];
          *reinterpret_cast<bf16x8*>(&stB[at]) = dz1f[j][rnd];
          *reinterpret_cast<bf16x8*>(&stC[at]) = dz0f[j][rnd];
          *reinterpret_cast<bf16x8*>(&stD[at]) = a1f[j][rnd];
        }
        *reinterpret_cast<uint2*>(&stY[row * 16 + 4 * g]) =
            make_uint2(pk_bf16(dy[rnd][0], dy[rnd][1]), pk_bf16(dy[rnd][2], dy[rnd][3]));
      }
      __syncthreads();
      // Every operand of the products below has the round's texels as its K
      // index: a fragment is two transposing reads of texel rows 32 kc + 8 g
      // + 4 h + k (k = 0..3 over the lanes 4 k + q), at columns c0 + 4 q.
      const int trr0 = 8 * g + (n16 >> 2), tq = 4 * (n16 & 3);
      auto frag_st = [&](const uint16_t* st, int stride, int kc, int c0) {
        const uint16_t* p = st + (32 * kc + trr0) * stride + c0 + tq;
        return frag_tr(p, p + 4 * stride);
      };
      // dW1[p_out][p_in] += dZ1 A0^T over positions; rows RW wave .. of this wave
      {
        bf16x8 a;
        lds_ring_ld<MTW * 2 * MT, kRing>(
            [&](int i) { return frag_st(stA, CF::SH, (i / MT) % 2, 16 * (i % MT)); },
            [&](int i, const bf16x8 b) {
              const int mt = i / (2 * MT), kc = (i / MT) % 2, c = i % MT;
              if (c == 0) a = frag_st(stB, CF::SH, kc, CF::RW * wave + 16 * mt);
              dw1[mt][c] = mfma(a, b, dw1[mt][c]);
            });
      }
      // dW0[p][f] += dZ0 X^T: the X^T fragments by transposing reads of the
      // X rows, K position p of the round at texel (p / 16) 16 NT + 16 rnd
      // + p % 16 (the staged texel order)
      {
        constexpr int NC = FP / 16;
        auto xrow = [&](int kc, int h) {
          const int p = 32 * kc + 8 * g + 4 * h + (n16 >> 2);
          return ((p >> 4) * 16 * kNT + 16 * rnd + (p & 15)) * CF::XS;
        };
        bf16x8 a;
        lds_ring_ld<MTW * 2 * NC, kRing>(
            [&](int i) {
              const int c = 16 * (i % NC) + tq, kc = (i / NC) % 2;
              return frag_tr(&xnf[xrow(kc, 0) + c], &xnf[xrow(kc, 1) + c]);
            },
            [&](int i, const bf16x8 b) {
              const int mt = i / (2 * NC), kc = (i / NC) % 2, c = i % NC;
              if (c == 0) a = frag_st(stC, CF::SH, kc, CF::RW * wave + 16 * mt);
              dw0[mt][c] = mfma(a, b, dw0[mt][c]);
            });
      }
      // dWout[o][p] += dY A1^T, column tiles of this wave
#pragma unroll
      for (int kc = 0; kc < 2; ++kc)
      {
        const bf16x8 a = frag_st(stY, 16, kc, 0);
#pragma unroll
        for (int c = 0; c < NTO; ++c) dwo[c] = mfma(a, frag_st(stD, CF::SH, kc, (wave * NTO + c) * 16), dwo[c]);
      }
      __syncthreads();
    }
    // the rounds ended with a barrier: the union region and the target words
    // are free for the next pass (the other table pair is unused); then the
    // latent gradients of this pass (its dX rows are complete since the box
    // barrier; the next pass rewrites them after its first barrier)
    fill_tables(nxt, atab + 2 * (tb ^ 1));
    prefetch_commit<CF, C0, C1>(pf, P, sbox);
    if (use_targets) targets_commit(tg, tgt_s, wave, n16, g);
    if (MODE != kForward && P.scatter && bx0.fits)
      box_gradients<C0, G0MODE == NTC_G0_CORNERS, CF::GS, kPass>(tab[0], bx0, dxs, 0, P.dg0, D0, wave, lane);
    if (MODE != kForward && P.scatter && bx1.fits)
    {
      // g1's tiles start on the wave after g0's last one
      const int t0 = bx0.fits ? ((bx0.w * bx0.h + 15) / 16) * ((C0 + 15) / 16) : 0;
      box_gradients<C1, false, CF::GS, kPass>(tab[1], bx1, dxs, CF::G1B, P.dg1, D1, (wave + 4 - (t0 & 3)) & 3,
                                              lane);
    }
    cur = nxt;
    tb ^= 1;
  }

  // ---- partial decoder gradients of this workgroup: W0[H][F] | W1[H][H] | Wout[16][H]
  if constexpr (MODE != kForward)
  {
  float* part = P.dw_part + (int64_t)vblk * P.dec_elems;
  const int F = P.F;
  float* pw0 = part;
  float* pw1 = part + H * F;
  float* pwo = pw1 + H * H;
#pragma unroll
  for (int mt = 0; mt < MTW; ++mt)
  {
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      // accumulator rows and dW1 columns are kperm positions
      const int row = kperm_pos(CF::RW * wave + 16 * mt + 4 * g + i);
#pragma unroll
      for (int c = 0; c < MT; ++c) pw1[row * H + kperm_pos(16 * c + n16)] = dw1[mt][c][i];
#pragma unroll
      for (int c = 0; c < FP / 16; ++c)
      {
        const int col = 16 * c + n16;
        if (col < F) pw0[row * F + col] = dw0[mt][c][i];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NTO; ++c)
#pragma unroll
    for (int i = 0; i < 4; ++i) pwo[(4 * g + i) * H + kperm_pos((wave * NTO + c) * 16 + n16)] = dwo[c][i];
  }

  // SSE of this workgroup (the union region is free after the last pass)
  __syncthreads();
  double* red = reinterpret_cast<double*>(uni);
  red[tid] = sse;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1)
  {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) P.sse_part[vblk] = red[0];
}


struct Shape
{
  int H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT;
};

// Column tiles per wave for a hidden width (register budget: H = 128 keeps
// one 16-texel tile per wave in flight).
constexpr int nt_for(int H) { return H >= 128 ? 1 : 2; }

int round_up(int n, int k) { return ((n + k - 1) / k) * k; }

bool shape_of(const ntc_feature_desc& fd, int H, int act, Shape& s)
{
  s.H = H;
  s.FP = round_up(fd.feature_dim, 32);
  const int g1b = fd.g0_mode == NTC_G0_CORNERS ? 4 * fd.g0_channels : fd.g0_channels;
  s.G = round_up(g1b + fd.g1_channels, 16);
  s.C0 = fd.g0_channels;
  s.C1 = fd.g1_channels;
  s.G0MODE = fd.g0_mode;
  s.PEC = fd.pe_count;
  s.LODS = fd.lod_slot;
  s.ACT = act;
  return true;
}


template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT>
constexpr Shape shape_v()
{
  return Shape{H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT};
}

// Layouts with a compiled kernel: the reference profiles (corner g0 features,
// HardGELU) and the measured configurations (bilinear g0, ReLU or HardGELU).
// X(...) has one tile shape, X2(...) also a 64-texel shape for the multi-set
// launch's shape tuning.
#define NTC_FUSED_LAYOUTS(X, X2)                                           \
  X(64, 64, 48, 8, 12, NTC_G0_CORNERS, 3, 1, NTC_ACT_HARDGELU)   /* Bpp_0_2 */  \
  X(64, 96, 80, 12, 20, NTC_G0_CORNERS, 3, 1, NTC_ACT_HARDGELU)  /* Bpp_0_5 */  \
  X(64, 96, 64, 12, 10, NTC_G0_CORNERS, 3, 1, NTC_ACT_HARDGELU)  /* Bpp_1_0 */  \
  X(64, 96, 80, 16, 12, NTC_G0_CORNERS, 3, 1, NTC_ACT_HARDGELU)  /* Bpp_2_25 */ \
  X2(64, 32, 32, 8, 16, NTC_G0_BILINEAR, 2, 0, NTC_ACT_RELU)     /* 8+16+8 -> 64 */ \
  X2(64, 64, 32, 12, 16, NTC_G0_BILINEAR, 2, 0, NTC_ACT_RELU)    /* 12+16+8 -> 64 */ \
  X(128, 64, 32, 16, 16, NTC_G0_BILINEAR, 2, 0, NTC_ACT_RELU)    /* 16+16+8 -> 128 */ \
  X2(64, 32, 32, 8, 16, NTC_G0_BILINEAR, 2, 0, NTC_ACT_HARDGELU)                \
  X(128, 64, 32, 16, 16, NTC_G0_BILINEAR, 2, 0, NTC_ACT_HARDGELU)

// The fused kernels map passes to texels with shifts: tile sides are powers of two.
bool pow2_tiles(const ntc_tiles& t) { return (t.tile_dim & (t.tile_dim - 1)) == 0; }

bool same(const Shape& a, const Shape& b)
{
  return a.H == b.H && a.FP == b.FP && a.G == b.G && a.C0 == b.C0 && a.C1 == b.C1 && a.G0MODE == b.G0MODE &&
         a.PEC == b.PEC && a.LODS == b.LODS && a.ACT == b.ACT;
}
} // namespace
