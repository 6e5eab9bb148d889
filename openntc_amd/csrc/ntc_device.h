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
    // each 32 x 32 -> 64-bit product as one v_mad_u64_u32 (the separate
    // low and high multiplies are two quarter-rate instructions)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
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
  // (x + 0.5) / M: for a power-of-two side (every mip side) the quotient is
  // exact, so one exponent shift gives the division's bits
  const T u = (M & (M - 1)) == 0 ? ldexp((T)x + (T)0.5, -__builtin_ctz((unsigned)M)) : ((T)x + (T)0.5) / (T)M;
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

// Bilinear corners of unwrapped texel coordinate pu (>= M after a wrap)
// along one axis, as unwrapped cell coordinates (cell + D after a wrap)
// minus o, and the fraction: the arithmetic of footprint().
__device__ __forceinline__ void axis_corners(int pu, int M, int D, int o, int& c0, int& c1, float& f)
{
  const bool wrapped = pu >= M;
  int x0;
  footprint_axis<float>(wrapped ? pu - M : pu, M, D, x0, f);
  const int add = (wrapped ? D : 0) - o;
  c0 = clampi(x0, 0, D - 1) + add;
  c1 = clampi(x0 + 1, 0, D - 1) + add;
}

// Cell rectangle (x0, len_x, y0, len_y), mod D, covered by the footprints of
// a tile's texels [ox, ox + T) x [oy, oy + T) (ox, oy in [0, M), T <= M).
__device__ __forceinline__ int4 tile_rect(int ox, int oy, int T, int M, int D)
{
  int a, b, c, d;
  float f;
  int4 r;
  axis_corners(ox, M, D, 0, a, b, f);
  axis_corners(ox + T - 1, M, D, 0, c, d, f);
  r.x = a;
  r.y = min(d - a + 1, D);
  axis_corners(oy, M, D, 0, a, b, f);
  axis_corners(oy + T - 1, M, D, 0, c, d, f);
  r.z = a;
  r.w = min(d - a + 1, D);
  return r;
}

// Whether cell (x, y) lies in one of the first n cell rectangles of a grid
// of side D.
__device__ __forceinline__ bool in_rects(const int4* r, int n, int x, int y, int D)
{
  bool in = false;
  for (int b = 0; b < n; ++b)
  {
    const int4 q = r[b];
    int dx = x - q.x, dy = y - q.z;
    dx += dx < 0 ? D : 0;
    dy += dy < 0 ? D : 0;
    in = in || (dx < q.y && dy < q.w);
  }
  return in;
}

// sse[blockIdx.x] += the sum of part[b * 16 + blockIdx.x] over b < n, one
// 256-thread block per channel: thread i sums b = i, i + 256, ... in order,
// then a fixed tree over the block. Deterministic for a given n.
__device__ __forceinline__ void sum_channel_partials(int n, const double* __restrict__ part,
                                                     double* __restrict__ sse)
{
  __shared__ double red[256];
  const int o = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += 256) s += part[(int64_t)b * 16 + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sse[o] += red[0];
}

// The device step counter of the multi-set launches is rewritten between
// launches; read it with a device-coherent vector load (not through the
// scalar cache, which later launches are not guaranteed to see refreshed)
// and broadcast it to the whole wave.
__device__ __forceinline__ int read_step(const int32_t* step)
{
  const int v = __hip_atomic_load(const_cast<int32_t*>(step), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_amdgcn_readfirstlane(v);
}

// Triangle-wave positional encoding value (ntc_kernel.cu:153-158).
template <typename T>
__device__ __forceinline__ T tri_wave(int coord, int P, int h)
{
  const T phase = h == 0 ? (T)0 : (T)0.25 * (T)P;
  const T t = (P & (P - 1)) == 0 ? ldexp((T)coord + (T)0.5 + phase, -__builtin_ctz((unsigned)P))
                                  : ((T)coord + (T)0.5 + phase) / (T)P;  // exact for power-of-two periods
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

// Adam constants of one update (bias1/bias2 are the reciprocal bias
// corrections 1 / (1 - beta^t)).
struct AdamConst
{
  float lr, b1, b2, omb1, omb2, bias1, bias2;
};

// One Adam step (ntc_kernel.cu:375-410):
// m = b1 m + (1-b1) g ; v = b2 v + ((1-b2) g) g ; p -= (lr (m bias1)) / (sqrt(v bias2) + 1e-8)
__device__ __forceinline__ void adam_one(const AdamConst& c, float g, float& m, float& v, float& p)
{
  m = fmaf(c.b1, m, c.omb1 * g);
  v = fmaf(c.b2, v, (c.omb2 * g) * g);
  p = p - (c.lr * (m * c.bias1)) / (__builtin_sqrtf(v * c.bias2) + 1e-8f);
}
// ---- gather (per texel) ------------------------------------------------------
// The features of texel t of a batch of tiles (ntc_kernel.cu:75-214):
// put(slot, value) for every slot < fd.feature_dim. Noise words are drawn per
// (corner, group of 4 channels); g1 (and g0 in bilinear mode) accumulate
// sum_c w_c*(G+N) with FMA in corner order.
template <typename T, class Put>
__device__ __forceinline__ void gather_texel(const ntc_feature_desc& fd, const ntc_tiles& tl, const ntc_noise& nz,
                                             const T* __restrict__ g0, const T* __restrict__ g1, int64_t t, Put&& put)
{
  const TexelPos p = texel_of(t, tl.origins, tl.tile_dim, tl.mip_dim);
  const Footprint<T> f0 = footprint<T>(p.x, p.y, tl.mip_dim, fd.g0_dim);
  const Footprint<T> f1 = footprint<T>(p.x, p.y, tl.mip_dim, fd.g1_dim);
  const int c0 = fd.g0_channels, c1 = fd.g1_channels;

  // g0
  if (fd.g0_mode == NTC_G0_CORNERS)
  {
    for (int c = 0; c < 4; ++c)
    {
      for (int q = 0; q * 4 < c0; ++q)
      {
        uint4 w = make_uint4(0, 0, 0, 0);
        if (nz.enabled) w = noise_words(nz.seed, nz.step, 0, nz.level, (uint32_t)f0.cell[c], q);
        for (int ch = 4 * q; ch < c0 && ch < 4 * q + 4; ++ch)
        {
          T v = g0[(int64_t)f0.cell[c] * c0 + ch];
          if (nz.enabled) v = v + (T)noise_value(pick(w, ch & 3), nz.delta0);
          put(c * c0 + ch, v);
        }
      }
    }
  }
  else
  {
    for (int q = 0; q * 4 < c0; ++q)
    {
      uint4 w[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        w[c] = nz.enabled ? noise_words(nz.seed, nz.step, 0, nz.level, (uint32_t)f0.cell[c], q)
                          : make_uint4(0, 0, 0, 0);
      for (int ch = 4 * q; ch < c0 && ch < 4 * q + 4; ++ch)
      {
        T acc = (T)0;
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
          T v = g0[(int64_t)f0.cell[c] * c0 + ch];
          if (nz.enabled) v = v + (T)noise_value(pick(w[c], ch & 3), nz.delta0);
          acc = fma(f0.w[c], v, acc);
        }
        put(ch, acc);
      }
    }
  }
  int slot = fd.g0_mode == NTC_G0_CORNERS ? 4 * c0 : c0;

  // g1, always bilinear
  for (int q = 0; q * 4 < c1; ++q)
  {
    uint4 w[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      w[c] = nz.enabled ? noise_words(nz.seed, nz.step, 1, nz.level, (uint32_t)f1.cell[c], q)
                        : make_uint4(0, 0, 0, 0);
    for (int ch = 4 * q; ch < c1 && ch < 4 * q + 4; ++ch)
    {
      T acc = (T)0;
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
        T v = g1[(int64_t)f1.cell[c] * c1 + ch];
        if (nz.enabled) v = v + (T)noise_value(pick(w[c], ch & 3), nz.delta1);
        acc = fma(f1.w[c], v, acc);
      }
      put(slot + ch, acc);
    }
  }
  slot += c1;

  for (int pi = 0; pi < fd.pe_count; ++pi)
  {
    const int P = fd.pe_first_period >> pi;
    for (int axis = 0; axis < 2; ++axis)
      for (int h = 0; h < 2; ++h)
      {
        put(slot, tri_wave<T>(axis == 0 ? p.x : p.y, P, h));
        ++slot;
      }
  }
  if (fd.lod_slot)
  {
    put(slot, (T)tl.norm_lod);
    ++slot;
  }
  for (; slot < fd.feature_dim; ++slot) put(slot, (T)0);
}

} // namespace ntc
