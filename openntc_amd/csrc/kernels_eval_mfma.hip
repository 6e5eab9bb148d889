// Evaluation decode on the matrix cores, at the rounding points of the VALU
// decode of kernels_eval.hip (whose fixed-order FMA chains stay the
// bit-exact reference) with the MFMA's FP32 accumulation order:
//   BF16: features, weights and hidden activations rounded to BF16 (nearest
//         even), v_mfma_f32_16x16x32_bf16;
//   FP8:  OCP E4M3 operands with per-layer scales (features x / s_x, weights
//         W / s_w, hidden activations a / s_a, saturating at 448),
//         v_mfma_f32_16x16x32_fp8_fp8 (exact products), each accumulator
//         scaled by s_w * s_in afterwards.
// Replaces, per evaluation batch, the reference's draw_features + three
// SGEMMs + activations + Sdot (libopenntc.cpp:972-1053).
//
// Each wave walks chunks of 64 texels on its own: every lane gathers one
// texel's features (noise off) into the wave's LDS rows in the operand format
// and its target values; then the wave decodes them, 16 x NTW texels per pass,
// with the weights as the MFMA A operand (the fused step's table layout: W0
// rows, W1 and Wout with permuted columns, so that a layer's two accumulator
// tiles are the next layer's B fragment), and adds the squared errors per
// output channel in FP64. Each workgroup writes its 16 channel sums in a
// fixed order.
#include <type_traits>

#include "eval_formats.h"
#include "fused_body.h"

namespace
{
constexpr int kHeaderBytes = 32;

// A texel's features with the layout known at compile time (C0 > 0): the
// gather of gather_texel (noise off) for power-of-two tile and mip sides,
// with the same corner weights and FMA order; channel quads of a corner cell
// as one 16-byte load when the channel count is a multiple of 4. C0 == 0:
// gather_texel itself (any layout).
template <int C0, int C1, int G0MODE, int PEC, int LODS, class Put>
__device__ __forceinline__ void gather_fixed(const ntc_feature_desc& fd, const ntc_tiles& tl, const uint32_t* orig,
                                             const float* g0, const float* g1, int64_t t, Put&& put)
{
  if constexpr (C0 == 0)
  {
    const ntc_noise nz = {};
    gather_texel<float>(fd, tl, nz, g0, g1, t, put);
  }
  else
  {
    const int T = tl.tile_dim, M = tl.mip_dim;
    const int lt = 31 - __builtin_clz((unsigned)T);
    const int b = (int)(t >> (2 * lt));
    const int xy = (int)(t & (((int64_t)1 << (2 * lt)) - 1));
    const uint32_t mm = (uint32_t)M - 1u;
    const int px = (int)(((uint32_t)(xy & (T - 1)) + (uint32_t)orig[2 * b]) & mm);
    const int py = (int)(((uint32_t)(xy >> lt) + (uint32_t)orig[2 * b + 1]) & mm);
    const Footprint<float> f0 = footprint<float>(px, py, M, fd.g0_dim);
    const Footprint<float> f1 = footprint<float>(px, py, M, fd.g1_dim);
    auto quads = [&](const float* grid, const Footprint<float>& f, int C, int q, float4 (&v)[4]) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
        const float* p = grid + (int64_t)f.cell[c] * C + 4 * q;
        if (C % 4 == 0)
          v[c] = *reinterpret_cast<const float4*>(p);
        else
          v[c] = make_float4(p[0], 4 * q + 1 < C ? p[1] : 0.0f, 4 * q + 2 < C ? p[2] : 0.0f, 4 * q + 3 < C ? p[3] : 0.0f);
      }
    };
    auto blend = [&](const float4 (&v)[4], const Footprint<float>& f, int base, int C, int q) {
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
        acc[0] = fmaf(f.w[c], v[c].x, acc[0]);
        acc[1] = fmaf(f.w[c], v[c].y, acc[1]);
        acc[2] = fmaf(f.w[c], v[c].z, acc[2]);
        acc[3] = fmaf(f.w[c], v[c].w, acc[3]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (4 * q + i < C) put(base + 4 * q + i, acc[i]);
    };
    if constexpr (G0MODE == NTC_G0_CORNERS)
    {
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int ch = 0; ch < C0; ++ch) put(c * C0 + ch, g0[(int64_t)f0.cell[c] * C0 + ch]);
    }
    else
    {
#pragma unroll
      for (int q = 0; q < (C0 + 3) / 4; ++q)
      {
        float4 v[4];
        quads(g0, f0, C0, q, v);
        blend(v, f0, 0, C0, q);
      }
    }
    constexpr int G1B = G0MODE == NTC_G0_CORNERS ? 4 * C0 : C0;
#pragma unroll
    for (int q = 0; q < (C1 + 3) / 4; ++q)
    {
      float4 v[4];
      quads(g1, f1, C1, q, v);
      blend(v, f1, G1B, C1, q);
    }
    int slot = G1B + C1;
    constexpr int PE_FIRST = PEC == 3 ? 8 : 4;
#pragma unroll
    for (int pi = 0; pi < PEC; ++pi)
#pragma unroll
      for (int axis = 0; axis < 2; ++axis)
#pragma unroll
        for (int h = 0; h < 2; ++h) put(slot++, tri_wave<float>(axis == 0 ? px : py, PE_FIRST >> pi, h));
    if (LODS) put(slot++, tl.norm_lod);
    for (; slot < fd.feature_dim; ++slot) put(slot, 0.0f);
  }
}

// Operand format of the matrix-core decode: BF16 (16-byte fragments of 8
// elements) or E4M3 (8-byte fragments).
template <bool FP8>
struct Op
{
  typedef typename std::conditional<FP8, uint8_t, uint16_t>::type El;
  typedef typename std::conditional<FP8, long, bf16x8>::type Frag;
  static constexpr int PAD = FP8 ? 16 : 8;  // elements after a texel row (16 bytes)
  __device__ static Frag load(const El* p)
  {
    if constexpr (FP8)
      return *reinterpret_cast<const long*>(p);
    else
      return frag_lds(p);
  }
  __device__ static f32x4 mma(const Frag& a, const Frag& b, const f32x4& c)
  {
    if constexpr (FP8)
      return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
    else
      return mfma(a, b, c);
  }
  // v * r clamped to the E4M3 range: v_cvt_pk_fp8_f32 (round to nearest
  // even) of it is the byte of fp8_q(v, r), which rounds to nearest even after
  // the same clamp
  __device__ static float sat(float v, float r) { return __builtin_amdgcn_fmed3f(v * r, -448.0f, 448.0f); }
  // operand element of v (FP8: the E4M3 byte of v * r)
  __device__ static uint32_t elem(float v, float r)
  {
    if constexpr (FP8)
      return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(sat(v, r), 0.0f, 0, false) & 0xFFu;
    else
      return bf16_dev(v);
  }
  // 8 values as one B fragment (FP8: of v * r)
  __device__ static Frag pack(const float (&v)[8], float r)
  {
    if constexpr (FP8)
    {
      uint32_t w[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
      {
        int x = __builtin_amdgcn_cvt_pk_fp8_f32(sat(v[4 * h], r), sat(v[4 * h + 1], r), 0, false);
        x = __builtin_amdgcn_cvt_pk_fp8_f32(sat(v[4 * h + 2], r), sat(v[4 * h + 3], r), x, true);
        w[h] = (uint32_t)x;
      }
      return (long)(((uint64_t)w[1] << 32) | w[0]);
    }
    else
      return frag_pack(v);
  }
};

template <bool FP8, int H, int FP, int NTW, bool GATHER, int ACT, int C0 = 0, int C1 = 0, int G0MODE = 0,
          int PEC = 0, int LODS = 0>
__global__ __launch_bounds__(kThreads) void k_decode_mfma(const uint8_t* __restrict__ table, int F, int O, int64_t B,
                                                        const float* __restrict__ x, const float* __restrict__ yt,
                                                        float* __restrict__ y, double* __restrict__ part,
                                                        ntc_feature_desc fd, ntc_tiles tl,
                                                        const float* __restrict__ g0, const float* __restrict__ g1,
                                                        const uint8_t* __restrict__ pkg)
{
  typedef Op<FP8> O8;
  typedef typename O8::El El;
  typedef typename O8::Frag Frag;
  constexpr int KF = FP / 32, KH = H / 32, MT = H / 16;
  constexpr int XS = FP + O8::PAD;  // elements per texel row
  constexpr int WE = H * FP + H * H + 16 * H;
  constexpr int SUB = 64 / (16 * NTW);  // decoder passes over a wave's 64 texels
  constexpr TableLayout TLY = {H, FP, 16};
  static_assert((WE * sizeof(El)) % 16 == 0 && (XS * sizeof(El)) % 16 == 0, "16-byte rows");
  __shared__ __attribute__((aligned(16))) El wt[WE];
  __shared__ __attribute__((aligned(16))) El xs_all[4 * 64 * XS];    // each wave's texel rows
  __shared__ __attribute__((aligned(16))) float ts_all[4 * 64 * 16];  // each wave's target values [texel][o]
  __shared__ float unorm[256];                                        // byte / 255, the division's bits
  __shared__ uint32_t orig_s[128];                                    // tile origins (gather: at most 64 tiles)
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, n16 = lane & 15, g = lane >> 4;
  // header: FP8 chain scales s_w0 s_x, s_w1 s_a0, s_wo s_a1 and the reciprocals r_x, r_a0, r_a1
  const float* hdr = reinterpret_cast<const float*>(table);
  const float sc0 = FP8 ? hdr[0] : 1.0f, sc1 = FP8 ? hdr[1] : 1.0f, sco = FP8 ? hdr[2] : 1.0f;
  const float rx = FP8 ? hdr[3] : 1.0f, ra0 = FP8 ? hdr[4] : 1.0f, ra1 = FP8 ? hdr[5] : 1.0f;
  {
    const uint4* src = reinterpret_cast<const uint4*>(table + kHeaderBytes);
    uint4* dst = reinterpret_cast<uint4*>(wt);
    for (int i = tid; i < (int)(WE * sizeof(El) / 16); i += kThreads) dst[i] = src[i];
    unorm[tid] = (float)tid / 255.0f;
    if (GATHER && tid < 2 * tl.n_tiles) orig_s[tid] = (uint32_t)tl.origins[tid];
  }
  __syncthreads();
  El* const xs = xs_all + wave * 64 * XS;
  float* const ts = ts_all + wave * 64 * 16;
  double sse[4] = {0.0, 0.0, 0.0, 0.0};
  // Each wave walks chunks of 64 texels on its own (its rows of xs and ts
  // are its own, so no workgroup barrier): every lane gathers one texel, then
  // the wave decodes the 64.
  for (int64_t c0 = ((int64_t)blockIdx.x * 4 + wave) * 64; c0 < B; c0 += (int64_t)gridDim.x * 4 * 64)
  {
    {
      const int64_t t = c0 + lane;
      const int64_t tc = t < B ? t : 0;
      // the texel's features in the operand format (slots arrive in order),
      // written as 32-bit words, zeros to FP
      uint32_t* row = reinterpret_cast<uint32_t*>(xs + lane * XS);
      constexpr int PER = 4 / sizeof(El);  // elements per word
      uint32_t acc = 0;
      auto put = [&](int slot, float v) {
        const int k = slot % PER;
        acc |= O8::elem(v, rx) << (8 * sizeof(El) * k);
        if (k == PER - 1)
        {
          row[slot / PER] = acc;
          acc = 0;
        }
      };
      if constexpr (GATHER)
      {
        int64_t tb;
        if constexpr (C0 == 0)
        {
          const TexelPos p = texel_of(tc, tl.origins, tl.tile_dim, tl.mip_dim);
          tb = ((int64_t)p.y * tl.mip_dim + p.x) * O;
        }
        else
        {
          // power-of-two sides: the texel position as shifts and masks
          const int T = tl.tile_dim, lt = 31 - __builtin_clz((unsigned)T);
          const int b = (int)(tc >> (2 * lt)), xy = (int)(tc & (((int64_t)1 << (2 * lt)) - 1));
          const uint32_t mm = (uint32_t)tl.mip_dim - 1u;
          const int px = (int)(((uint32_t)(xy & (T - 1)) + orig_s[2 * b]) & mm);
          const int py = (int)(((uint32_t)(xy >> lt) + orig_s[2 * b + 1]) & mm);
          tb = ((int64_t)py * tl.mip_dim + px) * O;
        }
        // the target bytes are loaded before the gather's loads so that all
        // of the texel's loads are in flight together
        uint32_t tg[16];
#pragma unroll
        for (int o = 0; o < 16; ++o) tg[o] = o < O ? pkg[tb + o] : 0u;
        gather_fixed<C0, C1, G0MODE, PEC, LODS>(fd, tl, orig_s, g0, g1, tc, put);
#pragma unroll
        for (int o = 0; o < 16; ++o)
          if (o < O) ts[lane * 16 + o] = unorm[tg[o]];
      }
      else
      {
        for (int k = 0; k < F; ++k) put(k, x[(int64_t)k * B + tc]);
        for (int o = 0; o < O; ++o) ts[lane * 16 + o] = yt[(int64_t)o * B + tc];
      }
      if (F % PER) row[F / PER] = acc;
      for (int k = (F + PER - 1) / PER; k < FP / PER; ++k) row[k] = 0;
    }
    // the wave's own LDS writes complete before its lanes read each other's rows
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int sb = 0; sb < SUB; ++sb)
    {
      const int tb0 = sb * 16 * NTW;  // first texel of this decoder pass within the wave's 64
      // layer 0: Z0 = W0 X, each weight fragment feeding NTW column tiles
      f32x4 z[MT][NTW];
#pragma unroll
      for (int r = 0; r < MT; ++r)
#pragma unroll
        for (int nt = 0; nt < NTW; ++nt) z[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < KF; ++j)
      {
        Frag xb[NTW];
#pragma unroll
        for (int nt = 0; nt < NTW; ++nt) xb[nt] = O8::load(&xs[(tb0 + nt * 16 + n16) * XS + 32 * j + 8 * g]);
#pragma unroll
        for (int r = 0; r < MT; ++r)
        {
          const int row = 16 * r + n16;
          const Frag a = O8::load(&wt[TLY.oW0() + row * FP + swz(row, 32 * j + 8 * g, FP)]);
#pragma unroll
          for (int nt = 0; nt < NTW; ++nt) z[r][nt] = O8::mma(a, xb[nt], z[r][nt]);
        }
      }
      // activation (of the scaled chain), the operand format, and two
      // accumulator tiles per B fragment of the next layer
      Frag af[KH][NTW];
      auto to_frags = [&](float sc, float r) {
#pragma unroll
        for (int j = 0; j < KH; ++j)
#pragma unroll
          for (int nt = 0; nt < NTW; ++nt)
          {
            float v[8];
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
              v[i] = act_fwd<float>(ACT, FP8 ? z[2 * j][nt][i] * sc : z[2 * j][nt][i]);
              v[4 + i] = act_fwd<float>(ACT, FP8 ? z[2 * j + 1][nt][i] * sc : z[2 * j + 1][nt][i]);
            }
            af[j][nt] = O8::pack(v, r);
          }
      };
      to_frags(sc0, ra0);
      // layer 1
#pragma unroll
      for (int r = 0; r < MT; ++r)
#pragma unroll
        for (int nt = 0; nt < NTW; ++nt) z[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < KH; ++j)
#pragma unroll
        for (int r = 0; r < MT; ++r)
        {
          const int row = 16 * r + n16;
          const Frag a = O8::load(&wt[TLY.oW1() + row * H + swz(row, 32 * j + 8 * g, H)]);
#pragma unroll
          for (int nt = 0; nt < NTW; ++nt) z[r][nt] = O8::mma(a, af[j][nt], z[r][nt]);
        }
      to_frags(sc1, ra1);
      // output layer and the squared errors of outputs 4g + i
#pragma unroll
      for (int nt = 0; nt < NTW; ++nt)
      {
        f32x4 yh = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < KH; ++j)
          yh = O8::mma(O8::load(&wt[TLY.oWo() + n16 * H + swz(n16, 32 * j + 8 * g, H)]), af[j][nt], yh);
        const int loc = tb0 + nt * 16 + n16;
        const int64_t t = c0 + loc;
        const bool live = t < B;
        const float4 tv = *reinterpret_cast<const float4*>(&ts[loc * 16 + 4 * g]);
        const float tgt[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
          const int o = 4 * g + i;
          const bool use = live && o < O;
          const float yo = FP8 ? yh[i] * sco : yh[i];
          const float r = tgt[i] - yo;
          sse[i] += use ? (double)r * (double)r : 0.0;
          if (!GATHER && y && use) y[(int64_t)o * B + t] = yo;
        }
      }
    }
    // the decoder passes' reads of xs and ts are done before the next gather writes them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // channel o = 4g + i: the 16 lanes of group g in every wave, in a fixed order
  __syncthreads();
  double* red = reinterpret_cast<double*>(xs_all);
  static_assert(4 * 64 * XS * sizeof(El) >= kThreads * 4 * 8, "the channel reduction reuses the feature rows");
#pragma unroll
  for (int i = 0; i < 4; ++i) red[tid * 4 + i] = sse[i];
  __syncthreads();
  if (tid < 16)
  {
    const int gg = tid >> 2, i = tid & 3;
    double s = 0.0;
    for (int w = 0; w < 4; ++w)
      for (int n = 0; n < 16; ++n) s += red[(w * 64 + 16 * gg + n) * 4 + i];
    part[(int64_t)blockIdx.x * 16 + tid] = s;
  }
}

__global__ __launch_bounds__(256) void k_decode_mfma_sum(int blocks, const double* __restrict__ part,
                                                         double* __restrict__ sse)
{
  sum_channel_partials(blocks, part, sse);
}

typedef void (*MfmaDecode)(const uint8_t*, int, int, int64_t, const float*, const float*, float*, double*,
                           ntc_feature_desc, ntc_tiles, const float*, const float*, const uint8_t*);

template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT>
constexpr bool layout_is(const ntc_feature_desc& fd, int h)
{
  return h == H && round_up(fd.feature_dim, 32) == FP && fd.g0_channels == C0 && fd.g1_channels == C1 &&
         fd.g0_mode == G0MODE && fd.pe_count == PEC && fd.lod_slot == LODS && fd.pe_first_period == (PEC == 3 ? 8 : 4);
}

#define NTC_EVAL_LAYOUT(H, FP, G, C0, C1, MODE, PEC, LODS, ACT)                                        \
  if (layout_is<H, FP, G, C0, C1, MODE, PEC, LODS, ACT>(fd, h) && act == ACT)                           \
    return k_decode_mfma<FP8, H, FP, H == 64 ? 4 : 2, true, ACT, C0, C1, MODE, PEC, LODS>;

// The gathering kernel specialised for a compiled layout and activation,
// else null.
template <bool FP8>
MfmaDecode mfma_kernel_fixed(const ntc_feature_desc& fd, int h, int act)
{
  NTC_FUSED_LAYOUTS(NTC_EVAL_LAYOUT, NTC_EVAL_LAYOUT)
  return nullptr;
}

template <bool FP8, bool GATHER, int ACT>
MfmaDecode mfma_kernel_act(int H, int FP)
{
  if (H == 64 && FP == 32) return k_decode_mfma<FP8, 64, 32, 4, GATHER, ACT>;
  if (H == 64 && FP == 64) return k_decode_mfma<FP8, 64, 64, 4, GATHER, ACT>;
  if (H == 64 && FP == 96) return k_decode_mfma<FP8, 64, 96, 4, GATHER, ACT>;
  if (H == 128 && FP == 32) return k_decode_mfma<FP8, 128, 32, 2, GATHER, ACT>;
  if (H == 128 && FP == 64) return k_decode_mfma<FP8, 128, 64, 2, GATHER, ACT>;
  if (H == 128 && FP == 96) return k_decode_mfma<FP8, 128, 96, 2, GATHER, ACT>;
  return nullptr;
}

template <bool FP8, bool GATHER>
MfmaDecode mfma_kernel(int H, int FP, int act)
{
  return act == NTC_ACT_RELU ? mfma_kernel_act<FP8, GATHER, NTC_ACT_RELU>(H, FP)
                             : mfma_kernel_act<FP8, GATHER, NTC_ACT_HARDGELU>(H, FP);
}

// Workgroups for a batch: persistent, two per CU (each copies the weight
// tables once and walks its share of the chunks), at most one per 256
// texels (the partial slots ntc_decode_workspace_bytes sizes).
int mfma_blocks(int64_t B)
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  const int64_t n = (B + 255) / 256;
  return (int)(n < 2 * (int64_t)cus ? n : 2 * (int64_t)cus);
}
} // namespace

// Host side of the table: a header of 8 floats (FP8: the chain scales
// s_w0 s_x, s_w1 s_a0, s_wo s_a1 and the reciprocals r_x, r_a0, r_a1), then
// the weights in the operand format, W0 [H][FP] | W1 [H][H] (columns
// permuted) | Wout [16][H] (columns permuted), in the fused step's layout,
// every row's 8-element chunks placed by the bank swizzle.
size_t ntc_mfma_table_bytes(bool fp8, int F, int H)
{
  return kHeaderBytes + (fp8 ? 1 : 2) * (size_t)(H * round_up(F, 32) + H * H + 16 * H);
}

// s: the FP8 scales {s_x, s_w0, s_w1, s_wo, s_a0, s_a1} (unused for BF16)
int ntc_mfma_table_build(bool fp8, int F, int H, int O, const float* dec, const float* s, void* table)
{
  const int FP = round_up(F, 32);
  std::memset(table, 0, kHeaderBytes);
  float* hdr = static_cast<float*>(table);
  float r[6] = {1, 1, 1, 1, 1, 1};
  if (fp8)
  {
    for (int i = 0; i < 6; ++i) r[i] = 1.0f / s[i];
    hdr[0] = s[1] * s[0];
    hdr[1] = s[2] * s[4];
    hdr[2] = s[3] * s[5];
    hdr[3] = r[0];
    hdr[4] = r[4];
    hdr[5] = r[5];
  }
  uint8_t* t8 = static_cast<uint8_t*>(table) + kHeaderBytes;
  uint16_t* t16 = reinterpret_cast<uint16_t*>(t8);
  auto put = [&](int idx, float v, int layer) {
    if (fp8)
      t8[idx] = e4m3_byte(fp8_q(v, r[layer]));
    else
      t16[idx] = f2bf(v);
  };
  const TableLayout L{H, FP, 16};
  const float* w0 = dec;
  const float* w1 = w0 + H * F;
  const float* wo = w1 + H * H;
  for (int m = 0; m < H; ++m)
    for (int c = 0; c < FP; ++c)
    {
      const int src = swz(m, c, FP);
      put(L.oW0() + m * FP + c, src < F ? w0[m * F + src] : 0.0f, 1);
    }
  for (int m = 0; m < H; ++m)
    for (int c = 0; c < H; ++c) put(L.oW1() + m * H + c, w1[m * H + kperm_pos(swz(m, c, H))], 2);
  for (int o = 0; o < 16; ++o)
    for (int c = 0; c < H; ++c) put(L.oWo() + o * H + c, o < O ? wo[o * H + kperm_pos(swz(o, c, H))] : 0.0f, 3);
  return NTC_OK;
}

int ntc_mfma_decode(bool fp8, bool gather, int H, int F, int O, int act, int64_t B, const void* table, const float* x,
                    const float* yt, float* y, double* sse, void* workspace, const ntc_feature_desc* fd,
                    const ntc_tiles* tl, const float* g0, const float* g1, const uint8_t* pkg, hipStream_t stream)
{
  const int FP = round_up(F, 32);
  NTC_RETURN_IF(act != NTC_ACT_RELU && act != NTC_ACT_HARDGELU, NTC_E_INVALID);
  MfmaDecode k = nullptr;
  if (gather) k = fp8 ? mfma_kernel_fixed<true>(*fd, H, act) : mfma_kernel_fixed<false>(*fd, H, act);
  if (!k)
    k = fp8 ? (gather ? mfma_kernel<true, true>(H, FP, act) : mfma_kernel<true, false>(H, FP, act))
            : (gather ? mfma_kernel<false, true>(H, FP, act) : mfma_kernel<false, false>(H, FP, act));
  NTC_RETURN_IF(!k, NTC_E_INVALID);
  NTC_RETURN_IF(gather && (!tl || !valid_tiles(tl) || !pow2_tiles(*tl)), NTC_E_INVALID);
  const int blocks = mfma_blocks(B);
  double* part = static_cast<double*>(workspace);
  const ntc_feature_desc fdv = fd ? *fd : ntc_feature_desc{};
  const ntc_tiles tlv = tl ? *tl : ntc_tiles{};
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, stream, static_cast<const uint8_t*>(table), F, O, B, x, yt, y,
                     part, fdv, tlv, g0, g1, pkg);
  hipLaunchKernelGGL(k_decode_mfma_sum, dim3(O), dim3(256), 0, stream, blocks, part, sse);
  return NTC_LAUNCH_STATUS();
}
