// Evaluation decode on the matrix cores: the BF16 decode as
// v_mfma_f32_16x16x32_bf16 chains with FP32 accumulation (features, weights
// and hidden activations rounded to BF16, nearest even, as in the VALU
// decode of kernels_eval.hip, whose fixed-order FMA chains stay the
// bit-exact reference). Replaces, per evaluation batch, the reference's
// draw_features + three SGEMMs + activations + Sdot (libopenntc.cpp:972-1053).
//
// A workgroup (4 waves) walks the batch in chunks of 4 x 16 x NTW texels:
// every thread gathers one texel's features (noise off) into an LDS row as
// BF16 and its target values; then each wave runs the decoder for its 16 x NTW
// texels with the weights as the MFMA A operand (the fused step's tables:
// W0 rows, W1 and Wout with permuted columns, so that a layer's two
// accumulator tiles are the next layer's B fragment) and adds the squared
// errors per output channel in FP64. Each workgroup writes its 16 channel sums
// in a fixed order.
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
__device__ __forceinline__ void gather_fixed(const ntc_feature_desc& fd, const ntc_tiles& tl, const float* g0,
                                             const float* g1, int64_t t, Put&& put)
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
    const int px = (int)(((uint32_t)(xy & (T - 1)) + (uint32_t)tl.origins[2 * b]) & mm);
    const int py = (int)(((uint32_t)(xy >> lt) + (uint32_t)tl.origins[2 * b + 1]) & mm);
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

template <int H, int FP, int NTW, bool GATHER, int ACT, int C0 = 0, int C1 = 0, int G0MODE = 0, int PEC = 0,
          int LODS = 0>
__global__ __launch_bounds__(kThreads) void k_decode_bf16_mfma(const uint16_t* __restrict__ wtab, int F, int O,
                                                             int64_t B, const float* __restrict__ x,
                                                             const float* __restrict__ yt, float* __restrict__ y,
                                                             double* __restrict__ part, ntc_feature_desc fd,
                                                             ntc_tiles tl, const float* __restrict__ g0,
                                                             const float* __restrict__ g1,
                                                             const uint8_t* __restrict__ pkg)
{
  constexpr int KF = FP / 32, KH = H / 32, MT = H / 16;
  constexpr int XS = FP + 8;  // BF16 elements per texel row (rows 16 bytes apart in bank groups)
  constexpr int WE = H * FP + H * H + 16 * H;
  constexpr int SUB = 64 / (16 * NTW);  // decoder passes over a wave's 64 texels
  constexpr TableLayout TLY = {H, FP, 16};
  __shared__ __attribute__((aligned(16))) uint16_t wt[WE];
  __shared__ __attribute__((aligned(16))) uint16_t xs_all[4 * 64 * XS];  // each wave's texel rows
  __shared__ __attribute__((aligned(16))) float ts_all[4 * 64 * 16];    // each wave's target values [texel][o]
  __shared__ float unorm[256];                                         // byte / 255, the division's bits
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, n16 = lane & 15, g = lane >> 4;
  {
    const uint4* src = reinterpret_cast<const uint4*>(wtab);
    uint4* dst = reinterpret_cast<uint4*>(wt);
    for (int i = tid; i < WE / 8; i += kThreads) dst[i] = src[i];
    unorm[tid] = (float)tid / 255.0f;
  }
  __syncthreads();
  uint16_t* const xs = xs_all + wave * 64 * XS;
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
      // the texel's features as BF16 pairs (slots arrive in order), zeros to FP
      uint32_t* row = reinterpret_cast<uint32_t*>(xs + lane * XS);
      uint32_t lo = 0;
      auto put = [&](int slot, float v) {
        if (slot & 1)
          row[slot >> 1] = lo | ((uint32_t)bf16_dev(v) << 16);
        else
          lo = bf16_dev(v);
      };
      if constexpr (GATHER)
      {
        gather_fixed<C0, C1, G0MODE, PEC, LODS>(fd, tl, g0, g1, tc, put);
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
          const int px = (int)(((uint32_t)(xy & (T - 1)) + (uint32_t)tl.origins[2 * b]) & mm);
          const int py = (int)(((uint32_t)(xy >> lt) + (uint32_t)tl.origins[2 * b + 1]) & mm);
          tb = ((int64_t)py * tl.mip_dim + px) * O;
        }
        for (int o = 0; o < O; ++o) ts[lane * 16 + o] = unorm[pkg[tb + o]];
      }
      else
      {
        for (int k = 0; k < F; ++k) put(k, x[(int64_t)k * B + tc]);
        for (int o = 0; o < O; ++o) ts[lane * 16 + o] = yt[(int64_t)o * B + tc];
      }
      if (F & 1) row[F >> 1] = lo;
      for (int k = (F + 1) >> 1; k < FP / 2; ++k) row[k] = 0;
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
        bf16x8 xb[NTW];
#pragma unroll
        for (int nt = 0; nt < NTW; ++nt) xb[nt] = frag_lds(&xs[(tb0 + nt * 16 + n16) * XS + 32 * j + 8 * g]);
#pragma unroll
        for (int r = 0; r < MT; ++r)
        {
          const int row = 16 * r + n16;
          const bf16x8 a = frag_lds(&wt[TLY.oW0() + row * FP + swz(row, 32 * j + 8 * g, FP)]);
#pragma unroll
          for (int nt = 0; nt < NTW; ++nt) z[r][nt] = mfma(a, xb[nt], z[r][nt]);
        }
      }
      // activation, BF16, and two accumulator tiles per B fragment of the next layer
      bf16x8 af[KH][NTW];
      auto to_frags = [&]() {
#pragma unroll
        for (int j = 0; j < KH; ++j)
#pragma unroll
          for (int nt = 0; nt < NTW; ++nt)
          {
            float v[8];
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
              v[i] = act_fwd<float>(ACT, z[2 * j][nt][i]);
              v[4 + i] = act_fwd<float>(ACT, z[2 * j + 1][nt][i]);
            }
            af[j][nt] = frag_pack(v);
          }
      };
      to_frags();
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
          const bf16x8 a = frag_lds(&wt[TLY.oW1() + row * H + swz(row, 32 * j + 8 * g, H)]);
#pragma unroll
          for (int nt = 0; nt < NTW; ++nt) z[r][nt] = mfma(a, af[j][nt], z[r][nt]);
        }
      to_frags();
      // output layer and the squared errors of outputs 4g + i
#pragma unroll
      for (int nt = 0; nt < NTW; ++nt)
      {
        f32x4 yh = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < KH; ++j)
          yh = mfma(frag_lds(&wt[TLY.oWo() + n16 * H + swz(n16, 32 * j + 8 * g, H)]), af[j][nt], yh);
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
          const float r = tgt[i] - yh[i];
          sse[i] += use ? (double)r * (double)r : 0.0;
          if (!GATHER && y && use) y[(int64_t)o * B + t] = yh[i];
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
  static_assert(4 * 64 * XS * 2 >= kThreads * 4 * 8, "the channel reduction reuses the feature rows");
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

typedef void (*MfmaDecode)(const uint16_t*, int, int, int64_t, const float*, const float*, float*, double*,
                           ntc_feature_desc, ntc_tiles, const float*, const float*, const uint8_t*);

template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT>
constexpr bool layout_is(const ntc_feature_desc& fd, int h)
{
  return h == H && round_up(fd.feature_dim, 32) == FP && fd.g0_channels == C0 && fd.g1_channels == C1 &&
         fd.g0_mode == G0MODE && fd.pe_count == PEC && fd.lod_slot == LODS && fd.pe_first_period == (PEC == 3 ? 8 : 4);
}

#define NTC_EVAL_LAYOUT(H, FP, G, C0, C1, MODE, PEC, LODS, ACT)                                        \
  if (layout_is<H, FP, G, C0, C1, MODE, PEC, LODS, ACT>(fd, h) && act == ACT)                           \
    return k_decode_bf16_mfma<H, FP, H == 64 ? 4 : 2, true, ACT, C0, C1, MODE, PEC, LODS>;

// The gathering kernel specialised for a compiled layout and activation,
// else null.
MfmaDecode mfma_kernel_fixed(const ntc_feature_desc& fd, int h, int act)
{
  NTC_FUSED_LAYOUTS(NTC_EVAL_LAYOUT, NTC_EVAL_LAYOUT)
  return nullptr;
}

template <bool GATHER, int ACT>
MfmaDecode mfma_kernel_act(int H, int FP)
{
  if (H == 64 && FP == 32) return k_decode_bf16_mfma<64, 32, 4, GATHER, ACT>;
  if (H == 64 && FP == 64) return k_decode_bf16_mfma<64, 64, 4, GATHER, ACT>;
  if (H == 64 && FP == 96) return k_decode_bf16_mfma<64, 96, 4, GATHER, ACT>;
  if (H == 128 && FP == 32) return k_decode_bf16_mfma<128, 32, 2, GATHER, ACT>;
  if (H == 128 && FP == 64) return k_decode_bf16_mfma<128, 64, 2, GATHER, ACT>;
  if (H == 128 && FP == 96) return k_decode_bf16_mfma<128, 96, 2, GATHER, ACT>;
  return nullptr;
}

template <bool GATHER>
MfmaDecode mfma_kernel(int H, int FP, int act)
{
  return act == NTC_ACT_RELU ? mfma_kernel_act<GATHER, NTC_ACT_RELU>(H, FP)
                             : mfma_kernel_act<GATHER, NTC_ACT_HARDGELU>(H, FP);
}

int chunk_of(int) { return 256; }

// Workgroups for a batch: persistent, two per CU (each copies the weight
// tables once and walks its share of the chunks), at most one per chunk and
// per 256 texels (the partial slots ntc_decode_workspace_bytes sizes).
int mfma_blocks(int64_t B, int H)
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  int64_t n = (B + chunk_of(H) - 1) / chunk_of(H);
  const int64_t cap = (B + 255) / 256;
  n = n < cap ? n : cap;
  return (int)(n < 2 * (int64_t)cus ? n : 2 * (int64_t)cus);
}
} // namespace

// Host side of the table: header, then the BF16 tables W0 [H][FP] | W1 [H][H]
// (columns permuted) | Wout [16][H] (columns permuted) in the fused step's
// layout, every row's 16-byte chunks placed by the bank swizzle.
size_t ntc_mfma_table_bytes(int F, int H) { return kHeaderBytes + 2 * (size_t)(H * round_up(F, 32) + H * H + 16 * H); }

int ntc_mfma_table_build(int F, int H, int O, const float* dec, void* table)
{
  const int FP = round_up(F, 32);
  std::memset(table, 0, kHeaderBytes);
  uint16_t* tab = reinterpret_cast<uint16_t*>(static_cast<uint8_t*>(table) + kHeaderBytes);
  const TableLayout L{H, FP, 16};
  const float* w0 = dec;
  const float* w1 = w0 + H * F;
  const float* wo = w1 + H * H;
  for (int m = 0; m < H; ++m)
    for (int c = 0; c < FP; ++c)
    {
      const int src = swz(m, c, FP);
      tab[L.oW0() + m * FP + c] = f2bf(src < F ? w0[m * F + src] : 0.0f);
    }
  for (int m = 0; m < H; ++m)
    for (int c = 0; c < H; ++c) tab[L.oW1() + m * H + c] = f2bf(w1[m * H + kperm_pos(swz(m, c, H))]);
  for (int o = 0; o < 16; ++o)
    for (int c = 0; c < H; ++c) tab[L.oWo() + o * H + c] = o < O ? f2bf(wo[o * H + kperm_pos(swz(o, c, H))]) : 0;
  return NTC_OK;
}

int ntc_mfma_decode(bool gather, int H, int F, int O, int act, int64_t B, const void* table, const float* x,
                    const float* yt, float* y, double* sse, void* workspace, const ntc_feature_desc* fd,
                    const ntc_tiles* tl, const float* g0, const float* g1, const uint8_t* pkg, hipStream_t stream)
{
  const int FP = round_up(F, 32);
  NTC_RETURN_IF(act != NTC_ACT_RELU && act != NTC_ACT_HARDGELU, NTC_E_INVALID);
  MfmaDecode k = gather ? mfma_kernel_fixed(*fd, H, act) : nullptr;
  if (!k) k = gather ? mfma_kernel<true>(H, FP, act) : mfma_kernel<false>(H, FP, act);
  NTC_RETURN_IF(!k, NTC_E_INVALID);
  NTC_RETURN_IF(gather && !pow2_tiles(*tl), NTC_E_INVALID);
  const int blocks = mfma_blocks(B, H);
  const uint16_t* w = reinterpret_cast<const uint16_t*>(static_cast<const uint8_t*>(table) + kHeaderBytes);
  double* part = static_cast<double*>(workspace);
  const ntc_feature_desc fdv = fd ? *fd : ntc_feature_desc{};
  const ntc_tiles tlv = tl ? *tl : ntc_tiles{};
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, stream, w, F, O, B, x, yt, y, part, fdv, tlv, g0, g1, pkg);
  hipLaunchKernelGGL(k_decode_mfma_sum, dim3(O), dim3(256), 0, stream, blocks, part, sse);
  return NTC_LAUNCH_STATUS();
}
