// Device-side building blocks shared by the gfx950 kernels: the Philox noise
// generator, the texel -> latent-grid geometry of the gather/scatter pair and
// the activations. Arithmetic is written operation by operation (the library
// is compiled with -ffp-contract=off) so results are reproducible bit for bit
// where the operation order is fixed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ntc_hip.h"

namespace ntc
{
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). Round:
// (hi0, lo0) = M0*c0, (hi1, lo1) = M1*c2, c = {hi1^c1^k0, lo1, hi0^c3^k1, lo0},
// key bumped by (W0, W1) between rounds.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r)
  {
    if (r > 0)
    {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c.x;
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

constexpr uint32_t kNoiseKey1 = 0x4E5443u;

// Four noise words for channels [4q, 4q+4) of one grid cell.
__device__ __forceinline__ uint4 noise_words(uint32_t seed, uint32_t step, int grid, int level,
                                             uint32_t cell, int q)
{
  return philox4x32_10(make_uint4(cell, (uint32_t)q, step, (uint32_t)grid | ((uint32_t)level << 1)),
                       seed, kNoiseKey1);
}

// U(0,1] * delta - delta/2, exact in float because delta is a power of two.
__device__ __forceinline__ float noise_value(uint32_t word, float delta)
{
  const float u = (float)((word >> 8) + 1u) * 0x1p-24f;
  return u * delta - 0.5f * delta;
}

__device__ __forceinline__ uint32_t pick(const uint4& w, int i)
{
  return i == 0 ? w.x : (i == 1 ? w.y : (i == 2 ? w.z : w.w));
}

__device__ __forceinline__ int clampi(int x, int a, int b) { return x > b ? b : (x < a ? a : x); }

// Texel position of sample t inside a batch of wrapping tiles
// (ntc_kernel.cu:94-99).
struct TexelPos
{
  int x, y, b, xy;
};

__device__ __forceinline__ TexelPos texel_of(int64_t t, const int32_t* origins, int T, int M)
{
  TexelPos p;
  const int TT = T * T;
  p.b = (int)(t / TT);
  p.xy = (int)(t - (int64_t)p.b * TT);
  // unsigned arithmetic keeps the position inside [0, M) even for an origin
  // outside [0, M); for valid origins it equals the signed form
  p.x = (int)(((uint32_t)(p.xy % T) + (uint32_t)origins[2 * p.b]) % (uint32_t)M);
  p.y = (int)(((uint32_t)(p.xy / T) + (uint32_t)origins[2 * p.b + 1]) % (uint32_t)M);
  return p;
}

// Bilinear footprint of a texel in a grid of side D (ntc_kernel.cu:100-115):
// corner c = j + 2k at (clamp(x0+j), clamp(y0+k)), weights
// {(1-fx)(1-fy), fx(1-fy), (1-fx)fy, fx fy}. All of it is exact for
// power-of-two M and D.
template <typename T>
struct Footprint
{
  int cell[4];
  T w[4];
};

// One axis of the footprint: unclamped first corner x0 and fraction f.
template <typename T>
__device__ __forceinline__ void footprint_axis(int x, int M, int D, int& x0, T& f)
{
  const T u = ((T)x + (T)0.5) / (T)M;
  const T g = u * (T)D - (T)0.5;
  const T fl = floor(g);
  x0 = (int)fl;
  f = g - fl;
}

template <typename T>
__device__ __forceinline__ Footprint<T> footprint(int x, int y, int M, int D)
{
  int x0, y0;
  T fx, fy;
  footprint_axis<T>(x, M, D, x0, fx);
  footprint_axis<T>(y, M, D, y0, fy);
  Footprint<T> f;
  const T one = (T)1;
  f.w[0] = (one - fx) * (one - fy);
  f.w[1] = fx * (one - fy);
  f.w[2] = (one - fx) * fy;
  f.w[3] = fx * fy;
#pragma unroll
  for (int c = 0; c < 4; ++c)
  {
    const int cx = clampi(x0 + (c & 1), 0, D - 1);
    const int cy = clampi(y0 + (c >> 1), 0, D - 1);
    f.cell[c] = cy * D + cx;
  }
  return f;
}

// Triangle-wave positional encoding value (ntc_kernel.cu:153-158).
template <typename T>
__device__ __forceinline__ T tri_wave(int coord, int P, int h)
{
  const T phase = h == 0 ? (T)0 : (T)0.25 * (T)P;
  const T t = ((T)coord + (T)0.5 + phase) / (T)P;
  const T s = t - floor(t);
  return (T)1 - (T)4 * fabs(s - (T)0.5);
}

// HardGELU (ntc_kernel.cu:13-29) and ReLU.
template <typename T>
__device__ __forceinline__ T act_fwd(int act, T z)
{
  if (act == NTC_ACT_HARDGELU)
    return z < (T)-1.5 ? (T)0 : (z < (T)1.5 ? (z / (T)3) * (z + (T)1.5) : z);
  return z > (T)0 ? z : (T)0;
}

template <typename T>
__device__ __forceinline__ T act_bwd(int act, T z, T g)
{
  if (act == NTC_ACT_HARDGELU)
    return z < (T)-1.5 ? (T)0 : (z < (T)1.5 ? (((T)2 * z + (T)1.5) * g) / (T)3 : g);
  return z > (T)0 ? g : (T)0;
}
} // namespace ntc
