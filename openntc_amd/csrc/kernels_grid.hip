// Optimiser and latent-grid maintenance kernels: Adam, the fused dense grid
// update, clamp, quantise, bit packing, max|x| and a graph-safe zero fill.
//
//   k_adam          ntc_kernel.cu:375-410 (update_adam)
//   k_grid_update   update_adam + clamp_grid (ntc_kernel.cu:468-482) + the
//                   next step's gradient FillZero (libopenntc.cpp:695-696),
//                   one HBM pass: reads p, m, v, g, writes p, m, v, g = 0.
//   k_grid_update_tiles  the same for a gradient that is zero outside the
//                   step's tile footprints: g is read and cleared only inside.
//   k_grid_lazy_catchup / k_grid_lazy_hot  the same update with the zero-
//                   gradient steps of cells outside the footprints deferred:
//                   each cell counts the steps applied to it and catches up
//                   (the same per-step arithmetic, in order) before anything
//                   reads it.
//   k_grid_update_cold / k_grid_update_hot  the two halves of that update:
//                   the cells outside every footprint rectangle (no gradient
//                   access; disjoint from what the step's training kernel
//                   reads or writes, so it may run beside it) and the cells
//                   inside (each once, enumerated per rectangle).
//   k_quantize      ntc_kernel.cu:448-466 (quantize_grid)
//   k_pack          ntc_kernel.cu:486-507 (quantize_pack)
//   k_max_abs       ntc_kernel.cu:511-531 (max_abs)
#include <cstring>

#include "ntc_device.h"
#include "ntc_host.h"

using namespace ntc;

namespace
{
constexpr int kThreads = 256;

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kThreads) void k_adam(int64_t n, AdamConst c, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   float* __restrict__ p)
{
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
  {
    float mi = m[i], vi = v[i], pi = p[i];
    adam_one(c, g[i], mi, vi, pi);
    m[i] = mi;
    v[i] = vi;
    p[i] = pi;
  }
}

// Vectorised: 4 params per thread per iteration with 16-byte loads/stores;
// streaming (non-temporal) hints since nothing is reused within the step.
__global__ __launch_bounds__(kThreads) void k_grid_update(int64_t n4, AdamConst c, float lo, float hi,
                                                          f4* __restrict__ g, f4* __restrict__ m,
                                                          f4* __restrict__ v, f4* __restrict__ p)
{
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * kThreads)
  {
    const f4 gi = __builtin_nontemporal_load(&g[i]);
    f4 mi = __builtin_nontemporal_load(&m[i]);
    f4 vi = __builtin_nontemporal_load(&v[i]);
    f4 pi = __builtin_nontemporal_load(&p[i]);
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
      float me = mi[e], ve = vi[e], pe = pi[e];
      adam_one(c, gi[e], me, ve, pe);
      mi[e] = me;
      vi[e] = ve;
      pi[e] = fmaxf(lo, fminf(hi, pe));
    }
    __builtin_nontemporal_store(mi, &m[i]);
    __builtin_nontemporal_store(vi, &v[i]);
    __builtin_nontemporal_store(pi, &p[i]);
    __builtin_nontemporal_store(zero, &g[i]);
  }
}

// Both grids of a level in one launch: items [0, a4) belong to grid 0, the
// rest to grid 1 (float4 items; both sizes are multiples of 4).
struct GridSeg
{
  f4* g;
  f4* m;
  f4* v;
  f4* p;
  float lo, hi;
};

__global__ __launch_bounds__(kThreads) void k_grid_update2(int64_t a4, int64_t n4, AdamConst c, GridSeg s0,
                                                           GridSeg s1)
{
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += (int64_t)gridDim.x * kThreads)
  {
    const bool second = i >= a4;
    const GridSeg& s = second ? s1 : s0;
    const int64_t k = second ? i - a4 : i;
    const f4 gi = __builtin_nontemporal_load(&s.g[k]);
    f4 mi = __builtin_nontemporal_load(&s.m[k]);
    f4 vi = __builtin_nontemporal_load(&s.v[k]);
    f4 pi = __builtin_nontemporal_load(&s.p[k]);
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
      float me = mi[e], ve = vi[e], pe = pi[e];
      adam_one(c, gi[e], me, ve, pe);
      mi[e] = me;
      vi[e] = ve;
      pi[e] = fmaxf(s.lo, fminf(s.hi, pe));
    }
    __builtin_nontemporal_store(mi, &s.m[k]);
    __builtin_nontemporal_store(vi, &s.v[k]);
    __builtin_nontemporal_store(pi, &s.p[k]);
    __builtin_nontemporal_store(zero, &s.g[k]);
  }
}

// The dense update for a gradient that is zero outside the tile footprints
// (one step's latent scatter): cells outside every footprint rectangle take
// the step with g = 0 without touching the gradient buffer. Channel counts
// are multiples of 4, so a float4 item never straddles two cells.
struct TileUpdate
{
  ntc_tiles tl;
  int D[2], Q[2];  // grid sides, channel quads
  int64_t a4, n4;  // float4 items of g0, of both grids
  GridSeg s[2];
  AdamConst c;
};

// One block's share (items vblk * 256 + tid, stepping vgrid * 256) of the
// update described by A.
__device__ __forceinline__ void grid_update_tiles_body(const TileUpdate& A, const int vblk, const int vgrid)
{
  __shared__ int4 rect[2][16];
  const int tid = threadIdx.x, n = A.tl.n_tiles;
  const int M = A.tl.mip_dim, T = A.tl.tile_dim;
  if (tid < 2 * n)
  {
    const int grid = tid >= n, b = grid ? tid - n : tid;
    const int ox = (int)((uint32_t)A.tl.origins[2 * b] % (uint32_t)M);
    const int oy = (int)((uint32_t)A.tl.origins[2 * b + 1] % (uint32_t)M);
    rect[grid][b] = tile_rect(ox, oy, T, M, A.D[grid]);
  }
  __syncthreads();
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int64_t i = (int64_t)vblk * kThreads + tid; i < A.n4; i += (int64_t)vgrid * kThreads)
  {
    const int grid = i >= A.a4;
    const GridSeg& s = A.s[grid];
    const int64_t k = grid ? i - A.a4 : i;
    const int D = A.D[grid];
    const uint32_t cell = (uint32_t)k / (uint32_t)A.Q[grid];
    const int y = (int)(cell / (uint32_t)D), x = (int)cell - y * D;
    const bool hot = in_rects(rect[grid], n, x, y, D);
    f4 mi = __builtin_nontemporal_load(&s.m[k]);
    f4 vi = __builtin_nontemporal_load(&s.v[k]);
    f4 pi = __builtin_nontemporal_load(&s.p[k]);
    const f4 gi = hot ? __builtin_nontemporal_load(&s.g[k]) : zero;
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
      float me = mi[e], ve = vi[e], pe = pi[e];
      adam_one(A.c, gi[e], me, ve, pe);
      mi[e] = me;
      vi[e] = ve;
      pi[e] = fmaxf(s.lo, fminf(s.hi, pe));
    }
    __builtin_nontemporal_store(mi, &s.m[k]);
    __builtin_nontemporal_store(vi, &s.v[k]);
    __builtin_nontemporal_store(pi, &s.p[k]);
    if (hot) __builtin_nontemporal_store(zero, &s.g[k]);
  }
}

__global__ __launch_bounds__(kThreads) void k_grid_update_tiles(TileUpdate A)
{
  grid_update_tiles_body(A, blockIdx.x, gridDim.x);
}

// Adam (g = 0 or the gradient item gi) + clamp of one float4 item.
__device__ __forceinline__ void tile_item_update(const TileUpdate& A, const GridSeg& s, int64_t k, const f4 gi)
{
  f4 mi = __builtin_nontemporal_load(&s.m[k]);
  f4 vi = __builtin_nontemporal_load(&s.v[k]);
  f4 pi = __builtin_nontemporal_load(&s.p[k]);
#pragma unroll
  for (int e = 0; e < 4; ++e)
  {
    float me = mi[e], ve = vi[e], pe = pi[e];
    adam_one(A.c, gi[e], me, ve, pe);
    mi[e] = me;
    vi[e] = ve;
    pi[e] = fmaxf(s.lo, fminf(s.hi, pe));
  }
  __builtin_nontemporal_store(mi, &s.m[k]);
  __builtin_nontemporal_store(vi, &s.v[k]);
  __builtin_nontemporal_store(pi, &s.p[k]);
}

__device__ __forceinline__ void load_rects(const TileUpdate& A, int4 (&rect)[2][16])
{
  const int tid = threadIdx.x, n = A.tl.n_tiles, M = A.tl.mip_dim, T = A.tl.tile_dim;
  if (tid < 2 * n)
  {
    const int grid = tid >= n, b = grid ? tid - n : tid;
    const int ox = (int)((uint32_t)A.tl.origins[2 * b] % (uint32_t)M);
    const int oy = (int)((uint32_t)A.tl.origins[2 * b + 1] % (uint32_t)M);
    rect[grid][b] = tile_rect(ox, oy, T, M, A.D[grid]);
  }
  __syncthreads();
}

// The items of grid_update_tiles_body outside every rectangle: the same
// arithmetic with g = 0; the gradient buffer is not touched. Grid sides are
// powers of two and the item -> cell division by Q <= 4 is a multiply-high
// by ceil(2^32 / Q) (exact below 2^31 items).
__global__ __launch_bounds__(kThreads) void k_grid_update_cold(TileUpdate A)
{
  __shared__ int4 rect[2][16];
  load_rects(A, rect);
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  const uint32_t qr[2] = {0xFFFFFFFFu / (uint32_t)A.Q[0] + 1u, 0xFFFFFFFFu / (uint32_t)A.Q[1] + 1u};
  const int lg[2] = {31 - __builtin_clz((unsigned)A.D[0]), 31 - __builtin_clz((unsigned)A.D[1])};
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < A.n4; i += (int64_t)gridDim.x * kThreads)
  {
    const int grid = i >= A.a4;
    const int64_t k = grid ? i - A.a4 : i;
    const int D = A.D[grid];
    // Q == 1: qr wraps to 0 and the cell is k itself
    const uint32_t cell = A.Q[grid] == 1 ? (uint32_t)k : __umulhi((uint32_t)k, qr[grid]);
    const int y = (int)(cell >> lg[grid]), x = (int)(cell & (uint32_t)(D - 1));
    if (in_rects(rect[grid], A.tl.n_tiles, x, y, D)) continue;
    tile_item_update(A, A.s[grid], k, zero);
  }
}

// ---- deferred zero-gradient steps ----
//
// last[grid][cell] counts the level's Adam steps applied to the cell (its
// p, m, v for every channel); tab[3 o + {0, 1, 2}] = lr, bias1, bias2 of the
// level's step o. A zero-gradient step is a function of (p, m, v) and the
// step's constants only, so applying steps last .. upto - 1 late, in order,
// gives the bits the eager update gives.
struct LazyArgs
{
  TileUpdate A;  // grids, Adam state, footprint tiles, clamp bounds; A.c: this step's constants (hot)
  const float* tab;
  float b1, b2;
  int32_t* last[2];
  int32_t upto;  // catch-up: apply steps last .. upto - 1; hot: this step's index
  ntc_tiles skip;  // catch-up of rectangles: cells in these tiles' rectangles are left out
  int32_t n_skip;  // 0: none
};

// Smallest v over the wave (every lane active), in a scalar register.
__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
  return __builtin_amdgcn_readfirstlane(v);
}

// Correctly rounded square root of y for y == 0 or 2^-96 <= y < +inf: the
// hardware estimate (within 1 ulp) and the choice between it and its two
// neighbours by the sign of the FMA residuals, the steps the compiler's exact
// sqrtf takes on that range (below 2^-96 it scales y by 2^32 first, and it
// passes 0 and +inf through; the residual test gives 0 for 0 by itself).
__device__ __forceinline__ float sqrt_rn_wide(float y)
{
  const float s = __builtin_amdgcn_sqrtf(y);
  const float sd = __int_as_float(__float_as_int(s) - 1), su = __int_as_float(__float_as_int(s) + 1);
  float r = fmaf(-sd, s, y) <= 0.0f ? sd : s;
  r = fmaf(-su, s, y) > 0.0f ? su : r;
  return r;
}

// Smallest / largest v over the wave (every lane active), wave-uniform.
__device__ __forceinline__ float wave_minf(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ float wave_maxf(float v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// a / b as the compiler's IEEE division computes it (reciprocal estimate,
// one Newton step, quotient and two residual corrections: Markstein's
// method) without its operand scaling (v_div_scale) and special-value
// fix-up (v_div_fixup): equal to a / b bit for bit where both are the
// identity, i.e. for a == 0 or 2^-90 <= |a|, b, |a / b| <= 2^60 (the scaling
// only acts near the ends of the exponent range, the fix-up only on zeros,
// infinities, NaNs and denormal or overflowing quotients). The residuals are
// formed negated (-fma(b, q, -a) = fma(-b, q, a) except for an exact zero,
// where it gives -0), so that a zero numerator keeps its sign as a / b does.
__device__ __forceinline__ float div_rn_mid(float a, float b)
{
  const float r = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, r, 1.0f);
  const float r2 = fmaf(e, r, r);
  const float q = a * r2;
  const float rem = -fmaf(b, q, -a);
  const float q2 = fmaf(rem, r2, q);
  const float rem2 = -fmaf(b, q2, -a);
  return fmaf(rem2, r2, q2);
}

// Zero-gradient steps lo .. upto - 1 of the lane's item (lo < upto, the
// wave's smallest counter), applied from the lane's own counter on where run.
// kCheckSqrt: some lane's y = v bias2 may fall below 2^-96 on the way (the
// exact sqrtf's scaled range); fast_lo, fast_hi: the step's quotients take
// div_rn_mid when lr is 0 or its bits lie in [fast_lo, fast_hi] (bounds over
// every running lane of the wave, from lazy_steps).
template <bool kCheckSqrt>
__device__ __forceinline__ void lazy_step_loop(const LazyArgs& L, const GridSeg& s, int from, bool run, int lo,
                                               uint32_t fast_lo, uint32_t fast_hi, f4& mi, f4& vi, f4& pi)
{
  AdamConst c;
  c.b1 = L.b1;
  c.b2 = L.b2;
  c.omb1 = 1.0f - L.b1;  // as make_adam
  c.omb2 = 1.0f - L.b2;
  // the step table is written before training and only read here: through
  // the constant address space the wave-uniform reads are scalar loads
  typedef __attribute__((address_space(4))) const float ConstFloat;
  ConstFloat* tab = (ConstFloat*)L.tab;
  float lr = tab[3 * lo], bias1 = tab[3 * lo + 1], bias2 = tab[3 * lo + 2];
  for (int o = lo; o < L.upto; ++o)
  {
    const int on = o + 1 < L.upto ? o + 1 : o;
    const float nlr = tab[3 * on], nb1 = tab[3 * on + 1], nb2 = tab[3 * on + 2];
    c.lr = lr;
    c.bias1 = bias1;
    c.bias2 = bias2;
    // wave-uniform (scalar) choice of the division for this step: the bit
    // patterns of non-negative floats order as the values do
    const uint32_t lrb = (uint32_t)__builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lr));
    const bool fast_div = lrb == 0u || (lrb >= fast_lo && lrb <= fast_hi);
    if (run && o >= from)
    {
      // adam_one with g = 0, its gradient terms (1-b1) 0 and ((1-b2) 0) 0
      // formed once (wave-uniform) instead of per element: the same values
      const float zm = c.omb1 * 0.0f, zv = (c.omb2 * 0.0f) * 0.0f;
      float me[4], y[4];
      bool low = false;
#pragma unroll
      for (int e = 0; e < 4; ++e)
      {
        me[e] = fmaf(c.b1, mi[e], zm);
        vi[e] = fmaf(c.b2, vi[e], zv);
        y[e] = vi[e] * c.bias2;
        if (kCheckSqrt) low |= y[e] > 0.0f && y[e] < 0x1p-96f;
      }
      // the exact sqrtf where some lane of the wave has a tiny positive y,
      // else the same bits without its scaling and special-value steps
      float r[4];
      if (kCheckSqrt && __builtin_amdgcn_ballot_w64(low) != 0)
      {
#pragma unroll
        for (int e = 0; e < 4; ++e) r[e] = __builtin_sqrtf(y[e]);
      }
      else
      {
#pragma unroll
        for (int e = 0; e < 4; ++e) r[e] = sqrt_rn_wide(y[e]);
      }
      float q[4];
      if (fast_div)
      {
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = div_rn_mid(c.lr * (me[e] * c.bias1), r[e] + 1e-8f);
      }
      else
      {
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = (c.lr * (me[e] * c.bias1)) / (r[e] + 1e-8f);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e)
      {
        const float pe = pi[e] - q[e];
        mi[e] = me[e];
        // fmaxf(lo, fminf(hi, pe)) for every non-NaN pe (lo <= hi)
        pi[e] = __builtin_amdgcn_fmed3f(pe, s.lo, s.hi);
      }
    }
    lr = nlr;
    bias1 = nb1;
    bias2 = nb2;
  }
}

// The step loop with the range checks it can leave out, from bounds of each
// running lane over its k = upto - from steps (factors of 2 of margin
// throughout; b1, b2 < 1 and every step's bias1 <= 1 / (1 - b1), bias2 <=
// 1 / (1 - b2), as for t >= 1):
//   * each step maps v to fl(b2 v), at least v b2' with b2' = b2 (1 - 2^-24),
//     and y = fl(v bias2) >= v (1 - 2^-24), so y >= 2^-96 on every step when
//     v >= 2^-95 b2'^-k (or v == 0: y stays 0, which the short sqrt gives);
//   * |m| falls by at most b1' = b1 (1 - 2^-24) a step and never grows, so
//     t = fl(m bias1) lies in [|m|min b1'^k / 2, 2 |m|max / (1 - b1)] and
//     b = sqrt(y) + 1e-8 in [1e-8, 2 sqrt(v_max / (1 - b2)) + 1e-7];
//     a = fl(lr t) and a / b stay in [2^-90, 2^60] for every lr in
//     [2^-90 max(1, b_max) / t_min, 2^60 / (1e8 t_max)] (a == 0 where m or lr
//     is zero).
__device__ __forceinline__ void lazy_steps(const LazyArgs& L, const GridSeg& s, int from, bool run, int lo, f4& mi,
                                           f4& vi, f4& pi)
{
  float tmin = __builtin_inff(), tmax = 0.0f, bmax = 0.0f;
  bool sqrt_ok = true;
  if (run)
  {
    const float k = (float)(L.upto - from);
    // per-step decay in binades, rounded up (v_log_f32 is within 1 ulp)
    const float d1 = -__builtin_amdgcn_logf(L.b1) * 1.001f + 1e-6f;
    const float d2 = -__builtin_amdgcn_logf(L.b2) * 1.001f + 1e-6f;
    float mmin = __builtin_inff(), mmax = 0.0f, vmin = __builtin_inff(), vmax = 0.0f;
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
      const float am = fabsf(mi[e]);
      if (am != 0.0f) mmin = fminf(mmin, am);
      mmax = fmaxf(mmax, am);
      if (vi[e] != 0.0f) vmin = fminf(vmin, vi[e]);
      vmax = fmaxf(vmax, vi[e]);
    }
    // log2 of the smallest v after k steps, one binade of margin
    sqrt_ok = vmin == __builtin_inff() || __builtin_amdgcn_logf(vmin) - k * d2 - 1.0f >= -95.0f;
    if (mmin != __builtin_inff()) tmin = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(mmin) - k * d1 - 2.0f);
    tmax = 2.0f * mmax / (1.0f - L.b1);
    bmax = 2.0f * __builtin_sqrtf(vmax / (1.0f - L.b2)) + 1e-7f;
  }
  const bool sqrt_ok_w = __builtin_amdgcn_ballot_w64(!sqrt_ok) == 0;
  tmin = wave_minf(tmin);
  tmax = wave_maxf(tmax);
  bmax = wave_maxf(bmax);
  // the lr range of the short division (empty: hi < lo); NaN bounds fail
  // every comparison below and leave it empty
  uint32_t fast_lo = 0xFFFFFFFFu, fast_hi = 0u;
  if (bmax <= 0x1p60f && tmin > 0.0f)
  {
    const float lr_lo = 0x1p-90f * fmaxf(1.0f, bmax) / tmin;  // 0 when every m is 0
    const float lr_hi = tmax > 0.0f ? 0x1p60f / (1e8f * tmax) : 0x1p60f;
    if (lr_lo <= lr_hi)
    {
      fast_lo = __builtin_bit_cast(uint32_t, lr_lo);
      fast_hi = __builtin_bit_cast(uint32_t, lr_hi);
    }
  }
  if (sqrt_ok_w)
    lazy_step_loop<false>(L, s, from, run, lo, fast_lo, fast_hi, mi, vi, pi);
  else
    lazy_step_loop<true>(L, s, from, run, lo, fast_lo, fast_hi, mi, vi, pi);
}

// Catch-up of one float4 item per lane to step L.upto (from: the item's
// cell counter; live: the lane holds an item). Called by every lane of the
// wave: the step loop runs from the wave's smallest counter with the step
// constants as wave-uniform loads (the next step's issued before the current
// one is applied), and a lane applies the steps from its own counter on.
// A lane whose m and v are all zero (a parameter no nonzero gradient has
// reached yet, so every cell before its first visit) owes only a clamp: each
// zero-gradient step keeps m = v = +0 and subtracts (lr (0 bias1)) / (0 +
// eps) = +0 from p, and the clamp is idempotent; it leaves the step loop to
// the other lanes.
__device__ __forceinline__ void lazy_run(const LazyArgs& L, const GridSeg& s, int64_t k, int from, bool live)
{
  const bool mine = live && from < L.upto;
  if (wave_min(mine ? from : L.upto) >= L.upto) return;
  f4 mi = {0.f, 0.f, 0.f, 0.f}, vi = mi, pi = mi;
  bool flat = false;
  if (mine)
  {
    mi = __builtin_nontemporal_load(&s.m[k]);
    vi = __builtin_nontemporal_load(&s.v[k]);
    pi = __builtin_nontemporal_load(&s.p[k]);
    flat = mi[0] == 0.0f && mi[1] == 0.0f && mi[2] == 0.0f && mi[3] == 0.0f && vi[0] == 0.0f && vi[1] == 0.0f &&
           vi[2] == 0.0f && vi[3] == 0.0f;
    if (flat)
    {
#pragma unroll
      for (int e = 0; e < 4; ++e)
      {
        mi[e] = 0.0f;
        vi[e] = 0.0f;
        pi[e] = __builtin_amdgcn_fmed3f(pi[e], s.lo, s.hi);
      }
    }
  }
  const bool run = mine && !flat;
  const int lo = wave_min(run ? from : L.upto);
  if (lo < L.upto)
    lazy_steps(L, s, from, run, lo, mi, vi, pi);
  if (mine)
  {
    __builtin_nontemporal_store(mi, &s.m[k]);
    __builtin_nontemporal_store(vi, &s.v[k]);
    __builtin_nontemporal_store(pi, &s.p[k]);
  }
}

// Catch-up of the cells inside the step's footprint rectangles (blockIdx.y =
// grid * 16 + rectangle, cells of earlier rectangles skipped) to step upto.
// Rectangle slot `slot` (grid * n_tiles + rectangle) of a rectangle walk,
// block bx of nbx on it.
__device__ __forceinline__ void lazy_catchup_rects_body(const LazyArgs& L, int bx, int nbx, int slot)
{
  __shared__ int4 rect[2][16], srect[2][16];
  const TileUpdate& A = L.A;
  if (threadIdx.x < 2 * L.n_skip)
  {
    const int n = L.n_skip, M = L.skip.mip_dim;
    const int g = threadIdx.x >= n, b = g ? threadIdx.x - n : threadIdx.x;
    const int ox = (int)((uint32_t)L.skip.origins[2 * b] % (uint32_t)M);
    const int oy = (int)((uint32_t)L.skip.origins[2 * b + 1] % (uint32_t)M);
    srect[g][b] = tile_rect(ox, oy, L.skip.tile_dim, M, A.D[g]);
  }
  load_rects(A, rect);  // ends with a barrier
  const int nt = A.tl.n_tiles;
  const int grid = slot >= nt, b = grid ? slot - nt : slot;
  const int4 r = rect[grid][b];
  const int D = A.D[grid], Q = A.Q[grid];
  // below 2^31 items (the grids' quads: checked on the host), 32-bit index
  // arithmetic. A wave takes cpw whole cells, Q items each (Q = 3 leaves a
  // lane idle), so each cell's counter is read and then written by one wave
  // and no other wave takes the cell.
  const uint32_t cells = (uint32_t)r.y * (uint32_t)r.w;
  const int cpw = 64 / Q;
  const int lane = (int)threadIdx.x & 63, lc = lane / Q, q = lane - lc * Q;
  const uint32_t per_block = (uint32_t)(kThreads / 64 * cpw);
  const uint32_t wave_first = (threadIdx.x >> 6) * (uint32_t)cpw;
  // uniform trip count: every lane of a wave takes part in lazy_run
  for (uint32_t c0 = (uint32_t)bx * per_block; c0 < cells; c0 += (uint32_t)nbx * per_block)
  {
    const uint32_t c = c0 + wave_first + (uint32_t)lc;
    bool live = lc < cpw && c < cells;
    int64_t k = 0, cell = 0;
    int from = L.upto;
    if (live)
    {
      const int iy = (int)(c / (uint32_t)r.y), ix = (int)c - iy * r.y;
      int x = r.x + ix, y = r.z + iy;
      x -= x >= D ? D : 0;
      y -= y >= D ? D : 0;
      live = !in_rects(rect[grid], b, x, y, D) && !in_rects(srect[grid], L.n_skip, x, y, D);
      cell = (int64_t)y * D + x;
      k = cell * Q + q;
      if (live) from = L.last[grid][cell];
    }
    lazy_run(L, A.s[grid], k, from, live);
    // every lane of the cell has read the counter (above, in this wave)
    if (live && q == 0 && from < L.upto) L.last[grid][cell] = L.upto;
  }
}

// Catch-up of the cells inside the step's footprint rectangles to step upto:
// blockIdx.y = rectangle slot, cells of earlier rectangles left out.
__global__ __launch_bounds__(kThreads) void k_grid_lazy_catchup_rects(LazyArgs L)
{
  if ((int)blockIdx.y >= 2 * L.A.tl.n_tiles) return;
  lazy_catchup_rects_body(L, blockIdx.x, gridDim.x, blockIdx.y);
}

// Catch-up of every cell of both grids to step upto.
__global__ __launch_bounds__(kThreads) void k_grid_lazy_catchup_all(LazyArgs L)
{
  const TileUpdate& A = L.A;
  // uniform trip count: every lane of a wave takes part in lazy_run (the
  // grid, g0 or g1, is per lane)
  for (int64_t i0 = (int64_t)blockIdx.x * kThreads; i0 < A.n4; i0 += (int64_t)gridDim.x * kThreads)
  {
    const int64_t i = i0 + threadIdx.x;
    const bool live = i < A.n4;
    const int grid = live && i >= A.a4;
    const int64_t k = !live ? 0 : (grid ? i - A.a4 : i);
    const int from = live ? L.last[grid][k / A.Q[grid]] : L.upto;
    lazy_run(L, A.s[grid], k, from, live);
  }
}

// The step with index upto on the cells inside the rectangles (as
// k_grid_update_hot, whose cells the catch-up made current), then
// last = upto + 1 for them.
__device__ __forceinline__ void lazy_hot_body(const LazyArgs& L, int bx, int nbx, int slot)
{
  __shared__ int4 rect[2][16];
  const TileUpdate& A = L.A;
  load_rects(A, rect);
  const int nt = A.tl.n_tiles;
  const int grid = slot >= nt, b = grid ? slot - nt : slot;
  const int4 r = rect[grid][b];
  const int D = A.D[grid], Q = A.Q[grid];
  const GridSeg& s = A.s[grid];
  const uint32_t items = (uint32_t)r.y * (uint32_t)r.w * (uint32_t)Q;  // below 2^31, as above
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t i = (uint32_t)bx * kThreads + threadIdx.x; i < items; i += (uint32_t)nbx * kThreads)
  {
    const uint32_t c = i / (uint32_t)Q;
    const int q = (int)(i - c * (uint32_t)Q);
    const int iy = (int)(c / (uint32_t)r.y), ix = (int)c - iy * r.y;
    int x = r.x + ix, y = r.z + iy;
    x -= x >= D ? D : 0;
    y -= y >= D ? D : 0;
    if (in_rects(rect[grid], b, x, y, D)) continue;
    const int64_t cell = (int64_t)y * D + x;
    const int64_t k = cell * Q + q;
    const f4 gi = __builtin_nontemporal_load(&s.g[k]);
    tile_item_update(A, s, k, gi);
    __builtin_nontemporal_store(zero, &s.g[k]);
    if (q == 0) L.last[grid][cell] = L.upto + 1;
  }
}

// The step with index upto on the cells inside the rectangles (as
// k_grid_update_hot, whose cells the catch-up made current), then
// last = upto + 1 for them; blockIdx.y = rectangle slot.
__global__ __launch_bounds__(kThreads) void k_grid_lazy_hot(LazyArgs L)
{
  if ((int)blockIdx.y >= 2 * L.A.tl.n_tiles) return;
  lazy_hot_body(L, blockIdx.x, gridDim.x, blockIdx.y);
}

// One launch for a step's hot part (blockIdx.y < 32: H, its rectangle slot)
// and the next step's catch-up (blockIdx.y >= 32: C, which leaves out H's
// rectangles when both are on one level), each on its own blocks: the two
// touch different cells. nbx_h / nbx_c: blocks per rectangle of each part.
__global__ __launch_bounds__(kThreads) void k_grid_lazy_hot_catchup(LazyArgs H, LazyArgs C, int nbx_h, int nbx_c)
{
  const int y = (int)blockIdx.y, bx = (int)blockIdx.x;
  if (y < 32)
  {
    if (y >= 2 * H.A.tl.n_tiles || bx >= nbx_h) return;
    lazy_hot_body(H, bx, nbx_h, y);
  }
  else
  {
    if (y - 32 >= 2 * C.A.tl.n_tiles || bx >= nbx_c) return;
    lazy_catchup_rects_body(C, bx, nbx_c, y - 32);
  }
}

__global__ __launch_bounds__(kThreads) void k_fill_i32(int64_t n, int32_t value, int32_t* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    out[i] = value;
}

// The items inside the rectangles: blockIdx.y = grid * 16 + rectangle b;
// a cell also inside an earlier rectangle of its grid is left to that one,
// so every item is updated once.
__global__ __launch_bounds__(kThreads) void k_grid_update_hot(TileUpdate A)
{
  __shared__ int4 rect[2][16];
  load_rects(A, rect);
  const int grid = blockIdx.y >> 4, b = blockIdx.y & 15;
  if (b >= A.tl.n_tiles) return;
  const int4 r = rect[grid][b];
  const int D = A.D[grid], Q = A.Q[grid];
  const GridSeg& s = A.s[grid];
  const int64_t items = (int64_t)r.y * r.w * Q;
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < items; i += (int64_t)gridDim.x * kThreads)
  {
    const int q = (int)(i % Q);
    const int c = (int)(i / Q);
    const int iy = c / r.y, ix = c - iy * r.y;
    int x = r.x + ix, y = r.z + iy;
    x -= x >= D ? D : 0;
    y -= y >= D ? D : 0;
    if (in_rects(rect[grid], b, x, y, D)) continue;
    const int64_t k = ((int64_t)y * D + x) * Q + q;
    const f4 gi = __builtin_nontemporal_load(&s.g[k]);
    tile_item_update(A, s, k, gi);
    __builtin_nontemporal_store(zero, &s.g[k]);
  }
}

// The grid updates of several sets for the step held in *step: set s owns
// blocks [blk_first[s], blk_first[s + 1]) (blk_first has n_sets + 1 entries
// per step; a set without a grid update this step owns none).
__global__ __launch_bounds__(kThreads) void k_grid_update_tiles_sets(const TileUpdate* __restrict__ recs,
                                                                     const int32_t* __restrict__ blk_first,
                                                                     int n_sets, const int32_t* __restrict__ step)
{
  const int t = read_step(step);
  const int32_t* first = blk_first + (int64_t)t * (n_sets + 1);
  const int b = blockIdx.x;
  if (b >= first[n_sets]) return;
  int s = 0;
  while (s + 1 < n_sets && b >= first[s + 1]) ++s;
  grid_update_tiles_body(recs[(int64_t)t * n_sets + s], b - first[s], first[s + 1] - first[s]);
}

// Deferred grid steps of several sets for the step held in *step: records
// [steps][n_sets] of LazyArgs, set s on x-blocks [blk_first[s],
// blk_first[s + 1]) of every rectangle slot blockIdx.y (blk_first as for
// k_grid_update_tiles_sets). HOT: the step's gradient step in each set's
// footprint cells; else the catch-up of those cells to the record's step.
template <bool HOT>
__global__ __launch_bounds__(kThreads) void k_grid_lazy_sets(const LazyArgs* __restrict__ recs,
                                                             const int32_t* __restrict__ blk_first, int n_sets,
                                                             const int32_t* __restrict__ step)
{
  const int t = read_step(step);
  const int32_t* first = blk_first + (int64_t)t * (n_sets + 1);
  const int b = blockIdx.x;
  if (b >= first[n_sets]) return;
  int s = 0;
  while (s + 1 < n_sets && b >= first[s + 1]) ++s;
  const LazyArgs& L = recs[(int64_t)t * n_sets + s];
  if ((int)blockIdx.y >= 2 * L.A.tl.n_tiles) return;
  if constexpr (HOT)
    lazy_hot_body(L, b - first[s], first[s + 1] - first[s], blockIdx.y);
  else
    lazy_catchup_rects_body(L, b - first[s], first[s + 1] - first[s], blockIdx.y);
}

__global__ __launch_bounds__(kThreads) void k_grid_update_tail(int64_t beg, int64_t n, AdamConst c,
                                                               float lo, float hi, float* __restrict__ g,
                                                               float* __restrict__ m, float* __restrict__ v,
                                                               float* __restrict__ p)
{
  const int64_t i = beg + (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float mi = m[i], vi = v[i], pi = p[i];
  adam_one(c, g[i], mi, vi, pi);
  m[i] = mi;
  v[i] = vi;
  p[i] = fmaxf(lo, fminf(hi, pi));
  g[i] = 0.0f;
}

__global__ __launch_bounds__(kThreads) void k_clamp(int64_t n, float lo, float hi, float* __restrict__ g)
{
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    g[i] = fmaxf(lo, fminf(hi, g[i]));
}

__global__ __launch_bounds__(kThreads) void k_quantize(int64_t n, float kmin, float kmax, float delta,
                                                       float* __restrict__ g)
{
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
  {
    const float k = __builtin_rintf(g[i] / delta);
    g[i] = fminf(fmaxf(k, kmin), kmax) * delta;
  }
}

// One 32-bit word per thread: codes of elements w*(32/b)+s at bits s*b.
__global__ __launch_bounds__(kThreads) void k_pack(int64_t words, int bits, const float* __restrict__ in,
                                                   uint32_t* __restrict__ out)
{
  const int64_t w = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (w >= words) return;
  const int per = 32 / bits;
  const int half = 1 << (bits - 1);
  uint32_t packed = 0;
  for (int s = 0; s < per; ++s)
  {
    int q = (int)__builtin_roundf(in[w * per + s] * (float)half);
    q = clampi(q, -half, half - 1);
    packed |= (uint32_t)(q + half) << (s * bits);
  }
  out[w] = packed;
}

__global__ __launch_bounds__(kThreads) void k_max_abs(int64_t n, const float* __restrict__ x,
                                                      float* __restrict__ result)
{
  __shared__ float red[kThreads];
  float mx = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    mx = fmaxf(mx, fabsf(x[i]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  // non-negative floats order like their bit patterns
  if (threadIdx.x == 0) atomicMax((int*)result, __float_as_int(red[0]));
}

__global__ __launch_bounds__(kThreads) void k_fill_zero(size_t n16, uint4* __restrict__ p16, size_t tail,
                                                        uint8_t* __restrict__ ptail)
{
  const uint4 z = make_uint4(0, 0, 0, 0);
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n16; i += (size_t)gridDim.x * kThreads)
    p16[i] = z;
  if (blockIdx.x == 0 && threadIdx.x < tail) ptail[threadIdx.x] = 0;
}

unsigned grid_stride_blocks(int64_t n)
{
  const int64_t b = (n + kThreads - 1) / kThreads;
  return (unsigned)(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

AdamConst make_adam(float lr, float b1, float b2, float bias1, float bias2)
{
  AdamConst c;
  c.lr = lr;
  c.b1 = b1;
  c.b2 = b2;
  c.omb1 = 1.0f - b1;
  c.omb2 = 1.0f - b2;
  c.bias1 = bias1;
  c.bias2 = bias2;
  return c;
}
} // namespace

extern "C" {
int ntc_adam_f32(int64_t n, float lr, float beta1, float beta2, float bias1, float bias2,
                 const float* grad, float* m, float* v, float* p, hipStream_t stream)
{
  NTC_RETURN_IF(n < 1 || !grad || !m || !v || !p, NTC_E_INVALID);
  hipLaunchKernelGGL(k_adam, dim3(grid_stride_blocks(n)), dim3(kThreads), 0, stream, n,
                     make_adam(lr, beta1, beta2, bias1, bias2), grad, m, v, p);
  return NTC_LAUNCH_STATUS();
}

int ntc_grid_update_f32(int64_t n, float lr, float beta1, float beta2, float bias1, float bias2,
                        float lo, float hi, float* grad, float* m, float* v, float* p,
                        hipStream_t stream)
{
  NTC_RETURN_IF(n < 1 || !grad || !m || !v || !p, NTC_E_INVALID);
  const AdamConst c = make_adam(lr, beta1, beta2, bias1, bias2);
  const bool aligned = (((uintptr_t)grad | (uintptr_t)m | (uintptr_t)v | (uintptr_t)p) & 15) == 0;
  const int64_t n4 = aligned ? n / 4 : 0;
  if (n4 > 0)
  {
    // 8 resident blocks per CU over 256 CUs; each thread then streams many float4s
    const int64_t want = (n4 + kThreads - 1) / kThreads;
    const unsigned blocks = (unsigned)(want < 2048 ? want : 2048);
    hipLaunchKernelGGL(k_grid_update, dim3(blocks), dim3(kThreads), 0, stream, n4, c, lo, hi,
                       (f4*)grad, (f4*)m, (f4*)v, (f4*)p);
  }
  const int64_t beg = n4 * 4;
  if (beg < n)
    hipLaunchKernelGGL(k_grid_update_tail, dim3(blocks_for(n - beg, kThreads)), dim3(kThreads), 0, stream,
                       beg, n, c, lo, hi, grad, m, v, p);
  return NTC_LAUNCH_STATUS();
}

int ntc_grid_update2_f32(int64_t n0, float* grad0, float* m0, float* v0, float* p0, float lo0, float hi0,
                         int64_t n1, float* grad1, float* m1, float* v1, float* p1, float lo1, float hi1, float lr,
                         float beta1, float beta2, float bias1, float bias2, hipStream_t stream)
{
  NTC_RETURN_IF(n0 < 4 || n1 < 4 || (n0 & 3) || (n1 & 3), NTC_E_INVALID);
  NTC_RETURN_IF(!grad0 || !m0 || !v0 || !p0 || !grad1 || !m1 || !v1 || !p1, NTC_E_INVALID);
  const uintptr_t bits = (uintptr_t)grad0 | (uintptr_t)m0 | (uintptr_t)v0 | (uintptr_t)p0 | (uintptr_t)grad1 |
                         (uintptr_t)m1 | (uintptr_t)v1 | (uintptr_t)p1;
  NTC_RETURN_IF(bits & 15, NTC_E_INVALID);
  const AdamConst c = make_adam(lr, beta1, beta2, bias1, bias2);
  const GridSeg s0 = {(f4*)grad0, (f4*)m0, (f4*)v0, (f4*)p0, lo0, hi0};
  const GridSeg s1 = {(f4*)grad1, (f4*)m1, (f4*)v1, (f4*)p1, lo1, hi1};
  const int64_t n4 = (n0 + n1) / 4;
  const int64_t want = (n4 + kThreads - 1) / kThreads;
  const unsigned blocks = (unsigned)(want < 2048 ? want : 2048);
  hipLaunchKernelGGL(k_grid_update2, dim3(blocks), dim3(kThreads), 0, stream, n0 / 4, n4, c, s0, s1);
  return NTC_LAUNCH_STATUS();
}

static int tile_update_of(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0, float* g1,
                          float* dg1, const ntc_grid_adam* adam, TileUpdate& A)
{
  NTC_RETURN_IF(!valid_feature_desc(fd) || !valid_tiles(tiles) || !adam, NTC_E_INVALID);
  NTC_RETURN_IF(tiles->n_tiles > 16 || fd->g0_channels % 4 || fd->g1_channels % 4, NTC_E_INVALID);
  NTC_RETURN_IF(!g0 || !dg0 || !g1 || !dg1 || !adam->m0 || !adam->v0 || !adam->m1 || !adam->v1, NTC_E_INVALID);
  const uintptr_t bits = (uintptr_t)g0 | (uintptr_t)dg0 | (uintptr_t)g1 | (uintptr_t)dg1 | (uintptr_t)adam->m0 |
                         (uintptr_t)adam->v0 | (uintptr_t)adam->m1 | (uintptr_t)adam->v1;
  NTC_RETURN_IF(bits & 15, NTC_E_INVALID);
  A.tl = *tiles;
  A.D[0] = fd->g0_dim;
  A.D[1] = fd->g1_dim;
  A.Q[0] = fd->g0_channels / 4;
  A.Q[1] = fd->g1_channels / 4;
  A.a4 = (int64_t)A.D[0] * A.D[0] * A.Q[0];
  A.n4 = A.a4 + (int64_t)A.D[1] * A.D[1] * A.Q[1];
  A.s[0] = GridSeg{(f4*)dg0, (f4*)adam->m0, (f4*)adam->v0, (f4*)g0, adam->lo0, adam->hi0};
  A.s[1] = GridSeg{(f4*)dg1, (f4*)adam->m1, (f4*)adam->v1, (f4*)g1, adam->lo1, adam->hi1};
  A.c = make_adam(adam->lr, adam->beta1, adam->beta2, adam->bias1, adam->bias2);
  return NTC_OK;
}

int ntc_grid_update_tiles_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0,
                              float* g1, float* dg1, const ntc_grid_adam* adam, hipStream_t stream)
{
  TileUpdate A;
  const int st = tile_update_of(fd, tiles, g0, dg0, g1, dg1, adam, A);
  if (st != NTC_OK) return st;
  const int64_t want = (A.n4 + kThreads - 1) / kThreads;
  const unsigned blocks = (unsigned)(want < 2048 ? want : 2048);
  hipLaunchKernelGGL(k_grid_update_tiles, dim3(blocks), dim3(kThreads), 0, stream, A);
  return NTC_LAUNCH_STATUS();
}

static int rect_blocks(const ntc_tiles* tiles, const TileUpdate& A, int64_t cap)
{
  // rectangle side bound: the corners of T texels at D / M cells per texel
  const int M = tiles->mip_dim, T = tiles->tile_dim;
  int64_t most = 0;
  for (int grid = 0; grid < 2; ++grid)
  {
    const int64_t D = A.D[grid];
    int64_t w = ((int64_t)(T - 1) * D + M - 1) / M + 3;
    w = w < D ? w : D;
    const int64_t items = w * w * A.Q[grid];
    most = items > most ? items : most;
  }
  const int64_t want = (most + kThreads - 1) / kThreads;
  return (int)(want < cap ? want : cap);
}

int ntc_grid_update_tiles_part_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0,
                                   float* g1, float* dg1, const ntc_grid_adam* adam, int32_t part, int32_t blocks,
                                   hipStream_t stream)
{
  NTC_RETURN_IF(part != NTC_TILES_COLD && part != NTC_TILES_HOT, NTC_E_INVALID);
  TileUpdate A;
  const int st = tile_update_of(fd, tiles, g0, dg0, g1, dg1, adam, A);
  if (st != NTC_OK) return st;
  // the cold kernel maps items to cells with shifts (power-of-two sides) and
  // a multiply-high by 1/Q (Q <= 4 channel quads)
  for (int grid = 0; grid < 2; ++grid)
    NTC_RETURN_IF((A.D[grid] & (A.D[grid] - 1)) || A.Q[grid] < 1 || A.Q[grid] > 4, NTC_E_INVALID);
  if (part == NTC_TILES_COLD)
  {
    const int64_t want = (A.n4 + kThreads - 1) / kThreads;
    const int64_t cap = blocks > 0 ? blocks : 2048;
    hipLaunchKernelGGL(k_grid_update_cold, dim3((unsigned)(want < cap ? want : cap)), dim3(kThreads), 0, stream, A);
  }
  else
    hipLaunchKernelGGL(k_grid_update_hot, dim3(rect_blocks(tiles, A, blocks > 0 ? blocks : 64), 32), dim3(kThreads), 0,
                       stream, A);
  return NTC_LAUNCH_STATUS();
}

static int lazy_args_of(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0, float* g1,
                        float* dg1, const ntc_grid_adam* adam, const float* tab, int32_t upto, int32_t* last0,
                        int32_t* last1, LazyArgs& L)
{
  NTC_RETURN_IF(!tab || !last0 || !last1 || upto < 0, NTC_E_INVALID);
  // the every-cell catch-up reads no tile: a one-texel placeholder passes the
  // checks (its origin pointer is never dereferenced)
  static const int32_t kNoOrigin[2] = {0, 0};
  ntc_tiles all = {kNoOrigin, 1, 1, 1, 0.0f};
  const int st = tile_update_of(fd, tiles ? tiles : &all, g0, dg0, g1, dg1, adam, L.A);
  if (st != NTC_OK) return st;
  // the rectangle walks index a grid's quads in 32 bits
  NTC_RETURN_IF(L.A.n4 >= ((int64_t)1 << 31), NTC_E_INVALID);
  L.tab = tab;
  L.b1 = adam->beta1;
  L.b2 = adam->beta2;
  L.last[0] = last0;
  L.last[1] = last1;
  L.upto = upto;
  L.skip = ntc_tiles{nullptr, 0, 1, 1, 0.0f};
  L.n_skip = 0;
  return NTC_OK;
}

int ntc_grid_lazy_catchup_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles, const ntc_tiles* skip,
                              float* g0, float* g1, const ntc_grid_adam* adam, const float* steps, int32_t upto,
                              const int32_t* last0, const int32_t* last1, hipStream_t stream)
{
  LazyArgs L;
  // the catch-up touches no gradient: any non-null pointer passes the checks
  const int st = lazy_args_of(fd, tiles, g0, g0, g1, g1, adam, steps, upto, const_cast<int32_t*>(last0),
                              const_cast<int32_t*>(last1), L);
  if (st != NTC_OK) return st;
  if (skip)
  {
    NTC_RETURN_IF(!tiles || !valid_tiles(skip) || skip->n_tiles > 16, NTC_E_INVALID);
    L.skip = *skip;
    L.n_skip = skip->n_tiles;
  }
  if (upto == 0) return NTC_OK;
  if (tiles)
    hipLaunchKernelGGL(k_grid_lazy_catchup_rects, dim3(rect_blocks(tiles, L.A, 256), 32), dim3(kThreads), 0, stream,
                       L);
  else
  {
    const int64_t want = (L.A.n4 + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_grid_lazy_catchup_all, dim3((unsigned)(want < 2048 ? want : 2048)), dim3(kThreads), 0,
                       stream, L);
  }
  return NTC_LAUNCH_STATUS();
}

int ntc_grid_lazy_hot_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0, float* g1,
                          float* dg1, const ntc_grid_adam* adam, int32_t step, int32_t* last0, int32_t* last1,
                          hipStream_t stream)
{
  NTC_RETURN_IF(!tiles, NTC_E_INVALID);
  LazyArgs L;
  static const float kUnused[3] = {0, 0, 0};
  const int st = lazy_args_of(fd, tiles, g0, dg0, g1, dg1, adam, kUnused, step, last0, last1, L);
  if (st != NTC_OK) return st;
  hipLaunchKernelGGL(k_grid_lazy_hot, dim3(rect_blocks(tiles, L.A, 256), 32), dim3(kThreads), 0, stream, L);
  return NTC_LAUNCH_STATUS();
}

int ntc_grid_lazy_hot_catchup_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0,
                                   float* g1, float* dg1, const ntc_grid_adam* adam, int32_t step, int32_t* last0,
                                   int32_t* last1, const ntc_feature_desc* next_fd, const ntc_tiles* next_tiles,
                                   const ntc_tiles* skip, float* next_g0, float* next_g1,
                                   const ntc_grid_adam* next_adam, const float* steps, int32_t upto,
                                   int32_t* next_last0, int32_t* next_last1, hipStream_t stream)
{
  NTC_RETURN_IF(!tiles || !next_tiles, NTC_E_INVALID);
  LazyArgs H, C;
  static const float kUnused[3] = {0, 0, 0};
  int st = lazy_args_of(fd, tiles, g0, dg0, g1, dg1, adam, kUnused, step, last0, last1, H);
  if (st != NTC_OK) return st;
  st = lazy_args_of(next_fd, next_tiles, next_g0, next_g0, next_g1, next_g1, next_adam, steps, upto, next_last0,
                    next_last1, C);
  if (st != NTC_OK) return st;
  if (skip)
  {
    NTC_RETURN_IF(!valid_tiles(skip) || skip->n_tiles > 16, NTC_E_INVALID);
    C.skip = *skip;
    C.n_skip = skip->n_tiles;
  }
  const int nbx_h = rect_blocks(tiles, H.A, 256);
  const int nbx_c = upto == 0 ? 0 : rect_blocks(next_tiles, C.A, 256);
  hipLaunchKernelGGL(k_grid_lazy_hot_catchup, dim3(nbx_h > nbx_c ? nbx_h : nbx_c, nbx_c > 0 ? 64 : 32),
                     dim3(kThreads), 0, stream, H, C, nbx_h, nbx_c);
  return NTC_LAUNCH_STATUS();
}

size_t ntc_grid_lazy_bytes(void) { return sizeof(LazyArgs); }

int ntc_grid_lazy_pack(const ntc_feature_desc* fd, const ntc_tiles* tiles, const ntc_tiles* skip, float* g0,
                       float* dg0, float* g1, float* dg1, const ntc_grid_adam* adam, const float* steps,
                       int32_t upto, int32_t* last0, int32_t* last1, void* record, int32_t* rect_blocks_out)
{
  NTC_RETURN_IF(!record || !tiles, NTC_E_INVALID);
  LazyArgs L;
  const int st = lazy_args_of(fd, tiles, g0, dg0, g1, dg1, adam, steps, upto, last0, last1, L);
  if (st != NTC_OK) return st;
  if (skip)
  {
    NTC_RETURN_IF(!valid_tiles(skip) || skip->n_tiles > 16, NTC_E_INVALID);
    L.skip = *skip;
    L.n_skip = skip->n_tiles;
  }
  std::memcpy(record, &L, sizeof(L));
  if (rect_blocks_out) *rect_blocks_out = rect_blocks(tiles, L.A, 256);
  return NTC_OK;
}

static int lazy_sets_launch(bool hot, const void* records, const int32_t* blk_first, int32_t n_sets,
                            const int32_t* step, int32_t grid, int32_t slots, hipStream_t stream)
{
  NTC_RETURN_IF(!records || !blk_first || !step || n_sets < 1 || grid < 1 || slots < 1 || slots > 32,
                NTC_E_INVALID);
  const LazyArgs* r = static_cast<const LazyArgs*>(records);
  if (hot)
    hipLaunchKernelGGL(k_grid_lazy_sets<true>, dim3(grid, slots), dim3(kThreads), 0, stream, r, blk_first, n_sets,
                       step);
  else
    hipLaunchKernelGGL(k_grid_lazy_sets<false>, dim3(grid, slots), dim3(kThreads), 0, stream, r, blk_first, n_sets,
                       step);
  return NTC_LAUNCH_STATUS();
}

int ntc_grid_lazy_catchup_sets(const void* records, const int32_t* blk_first, int32_t n_sets, const int32_t* step,
                               int32_t grid, int32_t slots, hipStream_t stream)
{
  return lazy_sets_launch(false, records, blk_first, n_sets, step, grid, slots, stream);
}

int ntc_grid_lazy_hot_sets(const void* records, const int32_t* blk_first, int32_t n_sets, const int32_t* step,
                           int32_t grid, int32_t slots, hipStream_t stream)
{
  return lazy_sets_launch(true, records, blk_first, n_sets, step, grid, slots, stream);
}

int ntc_fill_i32(int64_t n, int32_t value, int32_t* out, hipStream_t stream)
{
  NTC_RETURN_IF(n < 1 || !out, NTC_E_INVALID);
  const int64_t want = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)(want < 2048 ? want : 2048)), dim3(kThreads), 0, stream, n, value,
                     out);
  return NTC_LAUNCH_STATUS();
}

size_t ntc_grid_update_tiles_bytes(void) { return sizeof(TileUpdate); }

int ntc_grid_update_tiles_pack(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0, float* g1,
                               float* dg1, const ntc_grid_adam* adam, void* record, int64_t* items)
{
  NTC_RETURN_IF(!record, NTC_E_INVALID);
  TileUpdate A;
  const int st = tile_update_of(fd, tiles, g0, dg0, g1, dg1, adam, A);
  if (st != NTC_OK) return st;
  std::memcpy(record, &A, sizeof(A));
  if (items) *items = A.n4;
  return NTC_OK;
}

int ntc_grid_update_tiles_sets(const void* records, const int32_t* blk_first, int32_t n_sets, const int32_t* step,
                               int32_t grid, hipStream_t stream)
{
  NTC_RETURN_IF(!records || !blk_first || !step || n_sets < 1 || grid < 1, NTC_E_INVALID);
  hipLaunchKernelGGL(k_grid_update_tiles_sets, dim3(grid), dim3(kThreads), 0, stream,
                     static_cast<const TileUpdate*>(records), blk_first, n_sets, step);
  return NTC_LAUNCH_STATUS();
}

int ntc_clamp_grid_f32(int64_t n, float delta, float* g, hipStream_t stream)
{
  NTC_RETURN_IF(n < 1 || !g, NTC_E_INVALID);
  hipLaunchKernelGGL(k_clamp, dim3(grid_stride_blocks(n)), dim3(kThreads), 0, stream, n, -1.0f,
                     1.0f - delta, g);
  return NTC_LAUNCH_STATUS();
}

int ntc_quantize_grid_f32(int64_t n, int32_t bits, float delta, float* g, hipStream_t stream)
{
  NTC_RETURN_IF(n < 1 || !g || bits < 1 || bits > 8, NTC_E_INVALID);
  const float kmin = -(float)(1 << (bits - 1));
  const float kmax = (float)(1 << (bits - 1)) - 1.0f;
  hipLaunchKernelGGL(k_quantize, dim3(grid_stride_blocks(n)), dim3(kThreads), 0, stream, n, kmin, kmax,
                     delta, g);
  return NTC_LAUNCH_STATUS();
}

int ntc_quantize_pack_f32(int64_t n, int32_t bits, const float* in, uint32_t* out, hipStream_t stream)
{
  NTC_RETURN_IF(n < 1 || !in || !out, NTC_E_INVALID);
  NTC_RETURN_IF(bits != 1 && bits != 2 && bits != 4 && bits != 8, NTC_E_INVALID);
  NTC_RETURN_IF((n * bits) % 32 != 0, NTC_E_INVALID);
  const int64_t words = n * bits / 32;
  hipLaunchKernelGGL(k_pack, dim3(blocks_for(words, kThreads)), dim3(kThreads), 0, stream, words, bits, in,
                     out);
  return NTC_LAUNCH_STATUS();
}

int ntc_max_abs_f32(int64_t n, const float* x, float* result, hipStream_t stream)
{
  NTC_RETURN_IF(n < 1 || !x || !result, NTC_E_INVALID);
  hipLaunchKernelGGL(k_max_abs, dim3(grid_stride_blocks(n)), dim3(kThreads), 0, stream, n, x, result);
  return NTC_LAUNCH_STATUS();
}

int ntc_fill_zero(void* p, size_t bytes, hipStream_t stream)
{
  NTC_RETURN_IF(!p && bytes, NTC_E_INVALID);
  if (bytes == 0) return NTC_OK;
  const bool aligned = ((uintptr_t)p & 15) == 0;
  const size_t n16 = aligned ? bytes / 16 : 0;
  const size_t tail = bytes - n16 * 16;
  NTC_RETURN_IF(tail > kThreads, NTC_E_INVALID);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_stride_blocks((int64_t)(n16 ? n16 : 1))), dim3(kThreads), 0,
                     stream, n16, (uint4*)p, tail, (uint8_t*)p + n16 * 16);
  return NTC_LAUNCH_STATUS();
}
} // extern "C"
