/* ntc_oracle_decode.c - CPU restatement of the evaluation decode at reduced
 * precision. TEST INFRASTRUCTURE ONLY (see ntc_oracle.h).
 *
 * The forward of one texel (A.4: Z0 = W0 X, A0 = act(Z0), Z1 = W1 A0,
 * A1 = act(Z1), Y = Wout A1) evaluated four ways, each with a fixed operation
 * order so that the GPU decode can be compared bit for bit:
 *   FP32: one fmaf chain per output in ascending k from +0 (the reference's
 *         FP32 forward, libopenntc.cpp:651-655);
 *   BF16: the same chains over BF16-rounded weights, features and hidden
 *         activations (round to nearest even), FP32 accumulation;
 *   FP8:  OCP FP8 E4M3 operands with per-layer scales: features x / s_x with
 *         s_x = 1/256, each weight layer W / s_w with s_w = max|W| / 448,
 *         hidden activations a / s_a with s_a = max|a| / 448 over the whole
 *         evaluation; products are exact in FP32, each chain is scaled back
 *         by s_w * s_in after accumulation;
 *   INT8: the deployed decoder (openntc-gui/ntc_decode.hlsli:112-115,
 *         400-502): int8 features clamp(round(x * 128)), int32 dot products
 *         with the exported row-quantised weights, float(acc) * row scale,
 *         activation, * rcp_s_a, round, clamp to int8 per layer; output
 *         float(acc) * row scale. HLSL round() is DXIL Round_ne (ties to
 *         even). The bilinear g1 blend of the HLSL is taken as the training
 *         gather's fma chain (the HLSL mad contraction is hardware-defined).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "ntc_oracle.h"

static float o_act1(int32_t act, float v)
{
  if (act == 0) return v < -1.5f ? 0.0f : (v < 1.5f ? (v / 3.0f) * (v + 1.5f) : v);
  return v > 0.0f ? v : 0.0f;
}

static double o_act1d(int32_t act, double v)
{
  if (act == 0) return v < -1.5 ? 0.0 : (v < 1.5 ? (v / 3.0) * (v + 1.5) : v);
  return v > 0.0 ? v : 0.0;
}

static uint32_t o_bits(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

static float o_float(uint32_t u)
{
  float f;
  memcpy(&f, &u, 4);
  return f;
}

float o_bf16_round(float x)
{
  uint32_t u = o_bits(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return o_float(u & 0xFFFF0000u);
}

/* OCP FP8 E4M3: sign, 4 exponent bits (bias 7), 3 mantissa bits; largest
 * finite 448 (0x7E), 0x7F / 0xFF are NaN, no infinities. Rounds to nearest
 * even and saturates to +-448. Normal range: keep 3 of the 23 mantissa
 * bits with round-to-nearest-even on the float bits (a carry moves into the
 * exponent); below 2^-6: multiples of 2^-9 (subnormals, exact scaling). */
uint8_t o_e4m3_encode(float x)
{
  const uint32_t u = o_bits(x);
  const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
  float a = o_float(u & 0x7FFFFFFFu);
  if (!(a < 448.0f)) a = 448.0f; /* saturate (NaN inputs are not expected) */
  if (a < 0x1p-6f)
  {
    const float q = rintf(a * 512.0f); /* units of 2^-9, 0..8 */
    return (uint8_t)(sign | (uint8_t)q);
  }
  uint32_t v = o_bits(a);
  v += 0x7FFFFu + ((v >> 20) & 1u);
  const int32_t e = (int32_t)((v >> 23) & 0xFFu) - 127;
  const uint32_t m = (v >> 20) & 7u;
  if (e > 8 || (e == 8 && m == 7u)) return (uint8_t)(sign | 0x7Eu);
  return (uint8_t)(sign | (uint32_t)((e + 7) << 3) | m);
}

float o_e4m3_decode(uint8_t c)
{
  const int32_t e = (c >> 3) & 15, m = c & 7;
  float v = e == 0 ? (float)m * 0x1p-9f : ldexpf((float)(8 + m), e - 10);
  return (c & 0x80u) ? -v : v;
}

int32_t o_clamp_s8(float f)
{
  const float r = rintf(f);
  return r < -128.0f ? -128 : (r > 127.0f ? 127 : (int32_t)r);
}

static float o_maxabs(const float* w, int64_t n)
{
  float m = 0.0f;
  for (int64_t i = 0; i < n; ++i)
  {
    const float a = fabsf(w[i]);
    m = a > m ? a : m;
  }
  return m;
}

void o_fp8_scales(int32_t F, int32_t H, int32_t O, const float* dec, float max_a0, float max_a1, float s[6])
{
  const float* w0 = dec;
  const float* w1 = w0 + (int64_t)H * F;
  const float* wo = w1 + (int64_t)H * H;
  const float m0 = o_maxabs(w0, (int64_t)H * F), m1 = o_maxabs(w1, (int64_t)H * H);
  const float mo = o_maxabs(wo, (int64_t)O * H);
  s[0] = 1.0f / 256.0f;
  s[1] = m0 > 0.0f ? m0 / 448.0f : 1.0f;
  s[2] = m1 > 0.0f ? m1 / 448.0f : 1.0f;
  s[3] = mo > 0.0f ? mo / 448.0f : 1.0f;
  s[4] = max_a0 > 0.0f ? max_a0 / 448.0f : 1.0f;
  s[5] = max_a1 > 0.0f ? max_a1 / 448.0f : 1.0f;
}

static float o_q8(float v, float r) { return o_e4m3_decode(o_e4m3_encode(v * r)); }

void o_decode_f32(int32_t precision, int32_t act, int32_t F, int32_t Fp, int32_t H, int32_t O, int64_t B,
                  const float* dec, const float* s8, const uint8_t* blob, const float* rcp, const float* x,
                  float* y)
{
  const float* w0 = dec;
  const float* w1 = w0 ? w0 + (int64_t)H * F : NULL;
  const float* wo = w1 ? w1 + (int64_t)H * H : NULL;
  float r[6] = {1, 1, 1, 1, 1, 1};
  if (precision == 2)
    for (int i = 0; i < 6; ++i) r[i] = 1.0f / s8[i];
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < B; ++t)
  {
    float a0[256], a1[256];
    if (precision == 3)
    {
      const int8_t* q0 = (const int8_t*)blob;
      const int8_t* q1 = q0 + (int64_t)H * Fp;
      const int8_t* qo = q1 + (int64_t)H * H;
      const float* sc = (const float*)(qo + 16 * (int64_t)H);
      int32_t xi[256], h0[256], h1[256];
      for (int32_t k = 0; k < Fp; ++k) xi[k] = k < F ? o_clamp_s8(x[(int64_t)k * B + t] * 128.0f) : 0;
      for (int32_t h = 0; h < H; ++h)
      {
        int32_t acc = 0;
        for (int32_t k = 0; k < Fp; ++k) acc += (int32_t)q0[(int64_t)h * Fp + k] * xi[k];
        h0[h] = o_clamp_s8(o_act1(act, (float)acc * sc[h]) * rcp[0]);
      }
      for (int32_t h = 0; h < H; ++h)
      {
        int32_t acc = 0;
        for (int32_t k = 0; k < H; ++k) acc += (int32_t)q1[(int64_t)h * H + k] * h0[k];
        h1[h] = o_clamp_s8(o_act1(act, (float)acc * sc[H + h]) * rcp[1]);
      }
      for (int32_t o = 0; o < O; ++o)
      {
        int32_t acc = 0;
        for (int32_t k = 0; k < H; ++k) acc += (int32_t)qo[(int64_t)o * H + k] * h1[k];
        y[(int64_t)o * B + t] = (float)acc * sc[2 * H + o];
      }
      continue;
    }
    float xs[256];
    for (int32_t k = 0; k < F; ++k)
    {
      const float v = x[(int64_t)k * B + t];
      xs[k] = precision == 1 ? o_bf16_round(v) : (precision == 2 ? o_q8(v, r[0]) : v);
    }
    for (int32_t h = 0; h < H; ++h)
    {
      float acc = 0.0f;
      for (int32_t k = 0; k < F; ++k)
      {
        float w = w0[(int64_t)h * F + k];
        w = precision == 1 ? o_bf16_round(w) : (precision == 2 ? o_q8(w, r[1]) : w);
        acc = fmaf(w, xs[k], acc);
      }
      if (precision == 2) acc = acc * (s8[1] * s8[0]);
      const float a = o_act1(act, acc);
      a0[h] = precision == 1 ? o_bf16_round(a) : (precision == 2 ? o_q8(a, r[4]) : a);
    }
    for (int32_t h = 0; h < H; ++h)
    {
      float acc = 0.0f;
      for (int32_t k = 0; k < H; ++k)
      {
        float w = w1[(int64_t)h * H + k];
        w = precision == 1 ? o_bf16_round(w) : (precision == 2 ? o_q8(w, r[2]) : w);
        acc = fmaf(w, a0[k], acc);
      }
      if (precision == 2) acc = acc * (s8[2] * s8[4]);
      const float a = o_act1(act, acc);
      a1[h] = precision == 1 ? o_bf16_round(a) : (precision == 2 ? o_q8(a, r[5]) : a);
    }
    for (int32_t o = 0; o < O; ++o)
    {
      float acc = 0.0f;
      for (int32_t k = 0; k < H; ++k)
      {
        float w = wo[(int64_t)o * H + k];
        w = precision == 1 ? o_bf16_round(w) : (precision == 2 ? o_q8(w, r[3]) : w);
        acc = fmaf(w, a1[k], acc);
      }
      if (precision == 2) acc = acc * (s8[3] * s8[5]);
      y[(int64_t)o * B + t] = acc;
    }
  }
}

/* Blocks of NB texels side by side: each texel's chains keep the ascending k
 * order above, and the texel loop innermost lets the compiler vectorise the
 * fma calls across texels. */
#define O_NB 16
void o_decode_f64(int32_t act, int32_t F, int32_t H, int32_t O, int64_t B, const double* dec, const double* x,
                  double* y)
{
  const double* w0 = dec;
  const double* w1 = w0 + (int64_t)H * F;
  const double* wo = w1 + (int64_t)H * H;
  const int64_t nblk = (B + O_NB - 1) / O_NB;
#pragma omp parallel for schedule(static)
  for (int64_t blk = 0; blk < nblk; ++blk)
  {
    const int64_t t0 = blk * O_NB;
    const int32_t n = (int32_t)(B - t0 < O_NB ? B - t0 : O_NB);
    double xs[256][O_NB], a0[256][O_NB], a1[256][O_NB];
    for (int32_t k = 0; k < F; ++k)
      for (int32_t j = 0; j < O_NB; ++j) xs[k][j] = j < n ? x[(int64_t)k * B + t0 + j] : 0.0;
    for (int32_t h = 0; h < H; ++h)
    {
      double acc[O_NB] = {0};
      for (int32_t k = 0; k < F; ++k)
      {
        const double w = w0[(int64_t)h * F + k];
        for (int32_t j = 0; j < O_NB; ++j) acc[j] = fma(w, xs[k][j], acc[j]);
      }
      for (int32_t j = 0; j < O_NB; ++j) a0[h][j] = o_act1d(act, acc[j]);
    }
    for (int32_t h = 0; h < H; ++h)
    {
      double acc[O_NB] = {0};
      for (int32_t k = 0; k < H; ++k)
      {
        const double w = w1[(int64_t)h * H + k];
        for (int32_t j = 0; j < O_NB; ++j) acc[j] = fma(w, a0[k][j], acc[j]);
      }
      for (int32_t j = 0; j < O_NB; ++j) a1[h][j] = o_act1d(act, acc[j]);
    }
    for (int32_t o = 0; o < O; ++o)
    {
      double acc[O_NB] = {0};
      for (int32_t k = 0; k < H; ++k)
      {
        const double w = wo[(int64_t)o * H + k];
        for (int32_t j = 0; j < O_NB; ++j) acc[j] = fma(w, a1[k][j], acc[j]);
      }
      for (int32_t j = 0; j < n; ++j) y[(int64_t)o * B + t0 + j] = acc[j];
    }
  }
}
