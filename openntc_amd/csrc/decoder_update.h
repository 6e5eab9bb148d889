// The decoder update of one parameter (ordered sum of the workgroup
// partials + Adam + the BF16 table entry it changes), shared by the decoder
// update kernels and the training step's tail launch.
#pragma once

#include "fused_body.h"

namespace
{
// The table entries that pack_table_elem fills from decoder parameter e
// (W0[H][F] | W1[H][H] | Wout[16][H], row-major; a Wout row below O), set to
// b. The other entries (padding, Wout rows from O on) hold zeros that no
// update changes.
__host__ __device__ __forceinline__ void table_scatter(const TableLayout& L, int F, int gslots, int e, uint16_t b,
                                                       uint16_t* __restrict__ tab)
{
  const int H = L.H;
  if (e < H * F)
  {
    const int m = e / F, c = e % F;
    tab[L.oW0() + m * L.FP + swz(m, c, L.FP)] = b;
  }
  else if (e < H * F + H * H)
  {
    const int i = e - H * F, m = i / H, k = i % H;
    tab[L.oW1() + m * H + swz(m, kperm_col(k), H)] = b;
  }
  else
  {
    const int i = e - H * F - H * H, o = i / H, h = i % H;
    tab[L.oWo() + o * H + swz(o, kperm_col(h), H)] = b;
  }
}

// ---- decoder update: ordered sum of workgroup partials + Adam ----------------------
struct DecAdam
{
  float lr, b1, b2, omb1, omb2, bias1, bias2;
};

__device__ __forceinline__ void decoder_update_elem(int e, int n_dec, int n_adam_w01, int H, int O, int blocks,
                                                    const float* __restrict__ part, const DecAdam& c,
                                                    float* __restrict__ grad, float* __restrict__ m,
                                                    float* __restrict__ v, float* __restrict__ p,
                                                    const double* __restrict__ sse_part,
                                                    double* __restrict__ sse_out, const TableLayout& L = {},
                                                    int F = 0, int gslots = 0, uint16_t* __restrict__ tab = nullptr)
{
  if (e == 0 && sse_out)
  {
    // in block order, 16 loads in flight
    double s = 0.0;
    int b = 0;
    for (; b + 16 <= blocks; b += 16)
    {
      double w[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) w[k] = sse_part[b + k];
#pragma unroll
      for (int k = 0; k < 16; ++k) s += w[k];
    }
    for (; b < blocks; ++b) s += sse_part[b];
    *sse_out = s;
  }
  if (e >= n_dec) return;
  // Wout rows >= O are outside the optimised set (libopenntc.cpp:722-732)
  if (e >= n_adam_w01 && (e - n_adam_w01) / H >= O) return;
  // partials summed in block order; the loads of 32 partials are in flight
  // at once (the adds stay sequential)
  float gsum = 0.0f;
  int b = 0;
  for (; b + 32 <= blocks; b += 32)
  {
    float v[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = part[(int64_t)(b + k) * n_dec + e];
#pragma unroll
    for (int k = 0; k < 32; ++k) gsum = gsum + v[k];
  }
  for (; b < blocks; ++b) gsum = gsum + part[(int64_t)b * n_dec + e];
  grad[e] = gsum;
  const float mi = fmaf(c.b1, m[e], c.omb1 * gsum);
  const float vi = fmaf(c.b2, v[e], (c.omb2 * gsum) * gsum);
  m[e] = mi;
  v[e] = vi;
  const float pn = p[e] - (c.lr * (mi * c.bias1)) / (__builtin_sqrtf(vi * c.bias2) + 1e-8f);
  p[e] = pn;
  // the BF16 tables of the fused kernel follow the parameter (no separate pack)
  if (tab) table_scatter(L, F, gslots, e, f2bf(pn), tab);
}
} // namespace
