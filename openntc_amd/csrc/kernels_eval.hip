// Evaluation decode at reduced precision: the forward of every texel
// (Z0 = W0 X, A0 = act(Z0), Z1 = W1 A0, A1 = act(Z1), Y = Wout A1,
// libopenntc.cpp:651-655 / 972-1053) in FP32, BF16, FP8 (E4M3) or INT8, with
// the per-channel squared error against the targets summed in FP64.
//
// One thread per texel. Each output is one fixed-order chain: Z0[h] over the
// features in ascending order, Z1[h] over the hidden units, Y[o] over the
// hidden units as they are produced; so every precision is reproducible bit
// for bit on the CPU. The weight tables are read with uniform addresses
// (scalar loads) and built on the host from the FP32 decoder (FP32, BF16,
// FP8) or from the exported INT8 blob:
//   BF16: weights, features and hidden activations rounded to BF16 (nearest
//         even), FP32 accumulation;
//   FP8:  OCP E4M3 operands with per-layer scales (features x / s_x with
//         s_x = 1/256, weights W / s_w with s_w = max|W_layer| / 448, hidden
//         activations a / s_a with s_a = max|a| / 448 over the evaluation),
//         exact FP32 products, each chain scaled by s_w * s_in afterwards;
//   INT8: the deployed decoder's arithmetic (openntc-gui/ntc_decode.hlsli:
//         112-115, 400-502): int8 features clamp(round(x * 128)), int32 dot
//         products (v_dot4_i32_i8), float(acc) * row scale, activation,
//         * rcp_s_a, round to nearest even, clamp to int8; output
//         float(acc) * row scale.
#include <cstring>
#include <vector>

#include "eval_formats.h"
#include "ntc_device.h"
#include "ntc_host.h"

using namespace ntc;

// The matrix-core decode (kernels_eval_mfma.hip).
size_t ntc_mfma_table_bytes(bool fp8, int F, int H);
int ntc_mfma_table_build(bool fp8, int F, int H, int O, const float* dec, const float* s, void* table);
int ntc_mfma_decode(bool fp8, bool gather, int H, int F, int O, int act, int64_t B, const void* table, const float* x,
                    const float* yt, float* y, double* sse, void* workspace, const ntc_feature_desc* fd,
                    const ntc_tiles* tl, const float* g0, const float* g1, const uint8_t* pkg, hipStream_t stream);

namespace
{
constexpr int kThreads = 256;
constexpr int kHeader = 8;  // floats ahead of the weights in a table

__host__ __device__ inline int clamp_s8(float f)
{
  const float r = rintf(f);
  return r < -128.0f ? -128 : (r > 127.0f ? 127 : (int)r);
}

__device__ __forceinline__ float act1(int act, float z) { return act_fwd<float>(act, z); }

// ---- decode kernel -----------------------------------------------------------------------------
struct DecodeArgs
{
  int act, F, Fp, O;
  int64_t B;
  const float* table;
  const float* x;   // [F][B] (GATHER: unused)
  const float* yt;  // [O][B] (GATHER: unused)
  float* y;         // [O][B] or null
  double* part;     // [blocks][16]
  // GATHER: the batch's features and targets formed here (noise off)
  ntc_feature_desc fd;
  ntc_tiles tl;
  const float* g0;
  const float* g1;
  const uint8_t* pkg;  // the mip's targets [M][M][O]
};

// GATHER: each thread forms its texel's features (the gather of
// k_draw_features, the same function) into its column of the dynamic LDS
// array xs[F][256] and its targets from the package bytes (as
// k_draw_targets), so the batch's features never go through HBM; else X and
// the targets come from HBM.
template <int PREC, int H, bool GATHER>
__global__ __launch_bounds__(kThreads) void k_decode(DecodeArgs a)
{
  __shared__ double red[kThreads];
  extern __shared__ float xs[];
  const int tid = threadIdx.x;
  const int64_t t = (int64_t)blockIdx.x * kThreads + tid;
  const bool live = t < a.B;
  const int64_t tc = live ? t : 0;
  const float* __restrict__ hdr = a.table;
  if (GATHER)
  {
    const ntc_noise nz = {};
    gather_texel<float>(a.fd, a.tl, nz, a.g0, a.g1, tc, [&](int slot, float v) { xs[slot * kThreads + tid] = v; });
  }
  // feature k of this thread's texel
  auto xin = [&](int k) { return GATHER ? xs[k * kThreads + tid] : a.x[(int64_t)k * a.B + tc]; };
  float yv[16];
  if (PREC == NTC_PREC_INT8)
  {
    const uint32_t* __restrict__ q0t = reinterpret_cast<const uint32_t*>(hdr + kHeader);  // [Fp/4][H]
    const uint32_t* __restrict__ q1 = q0t + (a.Fp / 4) * H;                              // [H][H/4]
    const int* __restrict__ qot = reinterpret_cast<const int*>(q1 + H * (H / 4));         // [H][16]
    const float* __restrict__ isc = reinterpret_cast<const float*>(qot + H * 16);         // [2H + 16]
    const float rcp1 = hdr[0], rcp2 = hdr[1];
    int acc[H];
#pragma unroll
    for (int h = 0; h < H; ++h) acc[h] = 0;
    for (int k4 = 0; k4 < a.Fp / 4; ++k4)
    {
      uint32_t xp = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
        const int k = 4 * k4 + c;
        const int v = k < a.F ? clamp_s8(xin(k) * 128.0f) : 0;
        xp |= ((uint32_t)v & 0xFFu) << (8 * c);
      }
#pragma unroll
      for (int h = 0; h < H; ++h) acc[h] = __builtin_amdgcn_sdot4((int)q0t[k4 * H + h], (int)xp, acc[h], false);
    }
    uint32_t a0p[H / 4];
#pragma unroll
    for (int h4 = 0; h4 < H / 4; ++h4)
    {
      uint32_t p = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
        const int h = 4 * h4 + c;
        const int v = clamp_s8(act1(a.act, (float)acc[h] * isc[h]) * rcp1);
        p |= ((uint32_t)v & 0xFFu) << (8 * c);
      }
      a0p[h4] = p;
    }
    int yo[16];
#pragma unroll
    for (int o = 0; o < 16; ++o) yo[o] = 0;
    for (int h1 = 0; h1 < H; ++h1)
    {
      int s = 0;
#pragma unroll
      for (int j = 0; j < H / 4; ++j) s = __builtin_amdgcn_sdot4((int)q1[h1 * (H / 4) + j], (int)a0p[j], s, false);
      const int v = clamp_s8(act1(a.act, (float)s * isc[H + h1]) * rcp2);
#pragma unroll
      for (int o = 0; o < 16; ++o) yo[o] += qot[h1 * 16 + o] * v;
    }
#pragma unroll
    for (int o = 0; o < 16; ++o) yv[o] = (float)yo[o] * isc[2 * H + o];
  }
  else
  {
    const float* __restrict__ w0t = hdr + kHeader;      // [F][H]
    const float* __restrict__ w1 = w0t + a.F * H;       // [H][H]
    const float* __restrict__ wot = w1 + H * H;         // [H][16]
    const float sc0 = hdr[0], sc1 = hdr[1], sco = hdr[2], rx = hdr[3], ra0 = hdr[4], ra1 = hdr[5];
    float z[H];
#pragma unroll
    for (int h = 0; h < H; ++h) z[h] = 0.0f;
    for (int k = 0; k < a.F; ++k)
    {
      float xv = xin(k);
      xv = PREC == NTC_PREC_BF16 ? bf16_round(xv) : (PREC == NTC_PREC_FP8 ? fp8_q(xv, rx) : xv);
#pragma unroll
      for (int h = 0; h < H; ++h) z[h] = fmaf(w0t[k * H + h], xv, z[h]);
    }
#pragma unroll
    for (int h = 0; h < H; ++h)
    {
      const float v = act1(a.act, PREC == NTC_PREC_FP8 ? z[h] * sc0 : z[h]);
      z[h] = PREC == NTC_PREC_BF16 ? bf16_round(v) : (PREC == NTC_PREC_FP8 ? fp8_q(v, ra0) : v);
    }
#pragma unroll
    for (int o = 0; o < 16; ++o) yv[o] = 0.0f;
    for (int h1 = 0; h1 < H; ++h1)
    {
      float s = 0.0f;
#pragma unroll
      for (int h0 = 0; h0 < H; ++h0) s = fmaf(w1[h1 * H + h0], z[h0], s);
      float v = act1(a.act, PREC == NTC_PREC_FP8 ? s * sc1 : s);
      v = PREC == NTC_PREC_BF16 ? bf16_round(v) : (PREC == NTC_PREC_FP8 ? fp8_q(v, ra1) : v);
#pragma unroll
      for (int o = 0; o < 16; ++o) yv[o] = fmaf(wot[h1 * 16 + o], v, yv[o]);
    }
    if (PREC == NTC_PREC_FP8)
    {
#pragma unroll
      for (int o = 0; o < 16; ++o) yv[o] = yv[o] * sco;
    }
  }
  int64_t tbase = 0;
  if (GATHER)
  {
    const TexelPos p = texel_of(tc, a.tl.origins, a.tl.tile_dim, a.tl.mip_dim);
    tbase = ((int64_t)p.y * a.tl.mip_dim + p.x) * a.O;
  }
  // outputs and the per-channel squared error (fixed-order block sums)
  for (int o = 0; o < a.O; ++o)
  {
    double e = 0.0;
    if (live)
    {
      if (a.y) a.y[(int64_t)o * a.B + t] = yv[o];
      const float target = GATHER ? (float)a.pkg[tbase + o] / 255.0f : a.yt[(int64_t)o * a.B + t];
      const float r = target - yv[o];
      e = (double)r * (double)r;
    }
    red[tid] = e;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1)
    {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid == 0) a.part[(int64_t)blockIdx.x * 16 + o] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_decode_sum(int blocks, const double* __restrict__ part,
                                                    double* __restrict__ sse)
{
  sum_channel_partials(blocks, part, sse);
}

typedef void (*DecodeKernel)(DecodeArgs);

template <int H, bool GATHER = false>
DecodeKernel decode_kernel(int prec)
{
  switch (prec)
  {
  case NTC_PREC_FP32: return k_decode<NTC_PREC_FP32, H, GATHER>;
  case NTC_PREC_BF16: return k_decode<NTC_PREC_BF16, H, GATHER>;
  case NTC_PREC_FP8: return k_decode<NTC_PREC_FP8, H, GATHER>;
  case NTC_PREC_INT8: return k_decode<NTC_PREC_INT8, H, GATHER>;
  default: return nullptr;
  }
}

float max_abs(const float* w, int64_t n)
{
  float m = 0.0f;
  for (int64_t i = 0; i < n; ++i)
  {
    const float v = std::fabs(w[i]);
    m = v > m ? v : m;
  }
  return m;
}
} // namespace

extern "C" {
size_t ntc_decode_table_bytes(int32_t precision, int32_t feature_dim, int32_t feature_dim_padded, int32_t hidden_dim)
{
  const int64_t H = hidden_dim;
  if (precision == NTC_PREC_BF16_MFMA || precision == NTC_PREC_FP8_MFMA)
    return ntc_mfma_table_bytes(precision == NTC_PREC_FP8_MFMA, feature_dim, hidden_dim);
  if (precision == NTC_PREC_INT8)
    return sizeof(float) * kHeader + 4 * ((feature_dim_padded / 4) * H + H * (H / 4) + H * 16 + 2 * H + 16);
  return sizeof(float) * (kHeader + (int64_t)feature_dim * H + H * H + H * 16);
}

int ntc_build_decode_table_f(int32_t precision, int32_t F, int32_t H, int32_t O, const float* dec, float max_a0,
                             float max_a1, void* table)
{
  NTC_RETURN_IF(precision != NTC_PREC_FP32 && precision != NTC_PREC_BF16 && precision != NTC_PREC_FP8 &&
                    precision != NTC_PREC_BF16_MFMA && precision != NTC_PREC_FP8_MFMA,
                NTC_E_INVALID);
  NTC_RETURN_IF(!dec || !table || F < 1 || (H != 64 && H != 128) || O < 1 || O > 16, NTC_E_INVALID);
  const bool fp8 = precision == NTC_PREC_FP8 || precision == NTC_PREC_FP8_MFMA;
  const float* w0 = dec;
  const float* w1 = w0 + (int64_t)H * F;
  const float* wo = w1 + (int64_t)H * H;
  float* hdr = static_cast<float*>(table);
  float* w0t = hdr + kHeader;
  float* w1t = w0t + (int64_t)F * H;
  float* wot = w1t + (int64_t)H * H;
  // FP8 per-layer scales: s_x = 1/256, s_w = max|W| / 448, s_a = max|a| / 448
  float s[6] = {1.0f / 256.0f, 1, 1, 1, 1, 1};
  if (fp8)
  {
    const float m0 = max_abs(w0, (int64_t)H * F), m1 = max_abs(w1, (int64_t)H * H), mo = max_abs(wo, (int64_t)O * H);
    s[1] = m0 > 0.0f ? m0 / 448.0f : 1.0f;
    s[2] = m1 > 0.0f ? m1 / 448.0f : 1.0f;
    s[3] = mo > 0.0f ? mo / 448.0f : 1.0f;
    s[4] = max_a0 > 0.0f ? max_a0 / 448.0f : 1.0f;
    s[5] = max_a1 > 0.0f ? max_a1 / 448.0f : 1.0f;
  }
  if (precision == NTC_PREC_BF16_MFMA || precision == NTC_PREC_FP8_MFMA)
  {
    NTC_RETURN_IF(F > 96, NTC_E_INVALID);
    return ntc_mfma_table_build(fp8, F, H, O, dec, s, table);
  }
  float r[6];
  for (int i = 0; i < 6; ++i) r[i] = 1.0f / s[i];
  auto q = [&](float v, int layer) {
    return precision == NTC_PREC_BF16 ? bf16_round(v) : (precision == NTC_PREC_FP8 ? fp8_q(v, r[layer]) : v);
  };
  for (int h = 0; h < H; ++h)
    for (int k = 0; k < F; ++k) w0t[(int64_t)k * H + h] = q(w0[(int64_t)h * F + k], 1);
  for (int64_t i = 0; i < (int64_t)H * H; ++i) w1t[i] = q(w1[i], 2);
  for (int h = 0; h < H; ++h)
    for (int o = 0; o < 16; ++o) wot[h * 16 + o] = o < O ? q(wo[(int64_t)o * H + h], 3) : 0.0f;
  std::memset(hdr, 0, sizeof(float) * kHeader);
  if (precision == NTC_PREC_FP8)
  {
    hdr[0] = s[1] * s[0];
    hdr[1] = s[2] * s[4];
    hdr[2] = s[3] * s[5];
    hdr[3] = r[0];
    hdr[4] = r[4];
    hdr[5] = r[5];
  }
  return NTC_OK;
}

int ntc_build_decode_table_int8(int32_t F, int32_t Fp, int32_t H, int32_t O, const uint8_t* blob, float rcp_s_a1,
                                float rcp_s_a2, void* table)
{
  NTC_RETURN_IF(!blob || !table || F < 1 || Fp < F || Fp % 16 || (H != 64 && H != 128) || O < 1 || O > 16,
                NTC_E_INVALID);
  const int8_t* q0 = reinterpret_cast<const int8_t*>(blob);
  const int8_t* q1 = q0 + (int64_t)H * Fp;
  const int8_t* qo = q1 + (int64_t)H * H;
  const uint8_t* scb = reinterpret_cast<const uint8_t*>(qo + 16 * (int64_t)H);
  float* hdr = static_cast<float*>(table);
  std::memset(hdr, 0, sizeof(float) * kHeader);
  hdr[0] = rcp_s_a1;
  hdr[1] = rcp_s_a2;
  uint32_t* q0t = reinterpret_cast<uint32_t*>(hdr + kHeader);
  uint32_t* q1t = q0t + (Fp / 4) * H;
  int32_t* qot = reinterpret_cast<int32_t*>(q1t + (int64_t)H * (H / 4));
  float* isc = reinterpret_cast<float*>(qot + H * 16);
  auto pack4 = [](const int8_t* p) {
    uint32_t w = 0;
    for (int c = 0; c < 4; ++c) w |= ((uint32_t)(uint8_t)p[c]) << (8 * c);
    return w;
  };
  for (int k4 = 0; k4 < Fp / 4; ++k4)
    for (int h = 0; h < H; ++h) q0t[k4 * H + h] = pack4(q0 + (int64_t)h * Fp + 4 * k4);
  for (int h = 0; h < H; ++h)
    for (int j = 0; j < H / 4; ++j) q1t[h * (H / 4) + j] = pack4(q1 + (int64_t)h * H + 4 * j);
  for (int h = 0; h < H; ++h)
    for (int o = 0; o < 16; ++o) qot[h * 16 + o] = o < O ? (int32_t)qo[(int64_t)o * H + h] : 0;
  std::memcpy(isc, scb, sizeof(float) * (2 * H + 16));
  for (int o = O; o < 16; ++o) isc[2 * H + o] = 0.0f;
  return NTC_OK;
}

size_t ntc_decode_workspace_bytes(int64_t batch)
{
  return sizeof(double) * 16 * (size_t)blocks_for(batch < 1 ? 1 : batch, kThreads);
}

int ntc_decode_sse_f32(int32_t precision, int32_t activation, int32_t F, int32_t Fp, int32_t H, int32_t O,
                       int64_t batch, const void* table, const float* x, const float* y_target, double* sse,
                       float* y_out, void* workspace, hipStream_t stream)
{
  NTC_RETURN_IF(!table || !x || !y_target || !sse || !workspace || batch < 1, NTC_E_INVALID);
  NTC_RETURN_IF(F < 1 || F > 256 || O < 1 || O > 16 || (H != 64 && H != 128), NTC_E_INVALID);
  NTC_RETURN_IF(precision == NTC_PREC_INT8 && (Fp < F || Fp % 16), NTC_E_INVALID);
  if (precision == NTC_PREC_BF16_MFMA || precision == NTC_PREC_FP8_MFMA)
  {
    NTC_RETURN_IF(F > 96, NTC_E_INVALID);
    return ntc_mfma_decode(precision == NTC_PREC_FP8_MFMA, false, H, F, O, activation, batch, table, x, y_target,
                           y_out, sse, workspace, nullptr, nullptr, nullptr, nullptr, nullptr, stream);
  }
  const DecodeKernel k = H == 64 ? decode_kernel<64>(precision) : decode_kernel<128>(precision);
  NTC_RETURN_IF(!k, NTC_E_INVALID);
  DecodeArgs a;
  a.act = activation;
  a.F = F;
  a.Fp = Fp;
  a.O = O;
  a.B = batch;
  a.table = static_cast<const float*>(table);
  a.x = x;
  a.yt = y_target;
  a.y = y_out;
  a.part = static_cast<double*>(workspace);
  const int blocks = (int)blocks_for(batch, kThreads);
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, stream, a);
  hipLaunchKernelGGL(k_decode_sum, dim3(O), dim3(256), 0, stream, blocks, a.part, sse);
  return NTC_LAUNCH_STATUS();
}

int ntc_decode_tiles_sse_f32(int32_t precision, int32_t activation, const ntc_feature_desc* fd,
                             const ntc_tiles* tiles, const float* g0, const float* g1, const uint8_t* package,
                             int32_t Fp, int32_t H, int32_t O, const void* table, double* sse, void* workspace,
                             hipStream_t stream)
{
  NTC_RETURN_IF(!valid_feature_desc(fd) || !valid_tiles(tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!g0 || !g1 || !package || !table || !sse || !workspace, NTC_E_INVALID);
  const int F = fd->feature_dim;
  NTC_RETURN_IF(O < 1 || O > 16 || (H != 64 && H != 128), NTC_E_INVALID);
  NTC_RETURN_IF(precision == NTC_PREC_INT8 && (Fp < F || Fp % 16), NTC_E_INVALID);
  const int64_t batch = batch_of(*tiles);
  if (precision == NTC_PREC_BF16_MFMA || precision == NTC_PREC_FP8_MFMA)
  {
    NTC_RETURN_IF(F > 96, NTC_E_INVALID);
    return ntc_mfma_decode(precision == NTC_PREC_FP8_MFMA, true, H, F, O, activation, batch, table, nullptr,
                           nullptr, nullptr, sse, workspace, fd, tiles, g0, g1, package, stream);
  }
  const DecodeKernel k = H == 64 ? decode_kernel<64, true>(precision) : decode_kernel<128, true>(precision);
  NTC_RETURN_IF(!k, NTC_E_INVALID);
  DecodeArgs a = {};
  a.act = activation;
  a.F = F;
  a.Fp = Fp;
  a.O = O;
  a.B = batch;
  a.table = static_cast<const float*>(table);
  a.part = static_cast<double*>(workspace);
  a.fd = *fd;
  a.tl = *tiles;
  a.g0 = g0;
  a.g1 = g1;
  a.pkg = package;
  const int blocks = (int)blocks_for(batch, kThreads);
  const size_t lds = sizeof(float) * (size_t)F * kThreads;  // xs[F][256]
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), lds, stream, a);
  hipLaunchKernelGGL(k_decode_sum, dim3(O), dim3(256), 0, stream, blocks, a.part, sse);
  return NTC_LAUNCH_STATUS();
}
} // extern "C"
