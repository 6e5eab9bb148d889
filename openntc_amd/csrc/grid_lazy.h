// Device state and step loop of the deferred zero-gradient grid steps (the
// lazy Adam of the latent grids), shared by the grid kernels and the training
// step's tail launch, and the host helpers that describe one update.
#pragma once

#include "ntc_device.h"
#include "ntc_host.h"

using namespace ntc;

namespace
{
constexpr int kGridThreads = 256;  // threads per workgroup of every grid-update body

typedef float f4 __attribute__((ext_vector_type(4)));

// One grid of a level in an update: its gradient g, moments m, v and
// parameters p as float4 items, and its clamp bounds.
struct GridSeg
{
  f4* g;
  f4* m;
  f4* v;
  f4* p;
  float lo, hi;
};

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
  const uint32_t per_block = (uint32_t)(kGridThreads / 64 * cpw);
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

// ---- host side -------------------------------------------------------------

inline AdamConst make_adam(float lr, float b1, float b2, float bias1, float bias2)
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

static inline int tile_update_of(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0, float* g1,
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

static inline int rect_blocks(const ntc_tiles* tiles, const TileUpdate& A, int64_t cap)
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
  const int64_t want = (most + kGridThreads - 1) / kGridThreads;
  return (int)(want < cap ? want : cap);
}

static inline int lazy_args_of(const ntc_feature_desc* fd, const ntc_tiles* tiles, float* g0, float* dg0, float* g1,
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
} // namespace
