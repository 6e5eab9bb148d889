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
#include "grid_lazy.h"

using namespace ntc;

namespace
{
constexpr int kThreads = 256;

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
