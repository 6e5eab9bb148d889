// Plain path: one kernel per stage of the reference step, every activation
// materialised in HBM, FP32 (and FP64 for gradient checks). This is the
// unfused baseline the fused BF16 path is measured against, and the path that
// the CPU oracle checks stage by stage.
//
// Stage                          reference
//   k_draw_features              ntc_kernel.cu:75-214   (draw_features)
//   k_draw_targets               ntc_kernel.cu:216-259  (draw_targets)
//   k_gemm<AB|ATB|ABT>           libopenntc.cpp:395-417 (cuBLAS SGEMM wrappers)
//   k_act_fwd / k_act_bwd        ntc_kernel.cu:13-48    (hardgelu)
//   k_loss + k_sum_partials      libopenntc.cpp:659-668 (Saxpy, Sdot, scale)
//   k_scatter                    ntc_kernel.cu:261-373  (accumulate_grid_gradients)
#include "ntc_device.h"
#include "ntc_host.h"

using namespace ntc;

namespace
{
constexpr int kThreads = 256;

// ---- gather -----------------------------------------------------------------
// One thread per texel; writes X[f][t] (feature-major, coalesced across t).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_draw_features(ntc_feature_desc fd, ntc_tiles tl,
                                                            ntc_noise nz, const T* __restrict__ g0,
                                                            const T* __restrict__ g1,
                                                            T* __restrict__ x, int64_t B)
{
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= B) return;
  gather_texel<T>(fd, tl, nz, g0, g1, t, [&](int slot, T v) { x[(int64_t)slot * B + t] = v; });
}

// ---- targets ------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_draw_targets(ntc_tiles tl, int O,
                                                           const uint8_t* __restrict__ pkg,
                                                           T* __restrict__ y, int64_t B)
{
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= B) return;
  const TexelPos p = texel_of(t, tl.origins, tl.tile_dim, tl.mip_dim);
  const int64_t base = ((int64_t)p.y * tl.mip_dim + p.x) * O;
  for (int o = 0; o < O; ++o) y[(int64_t)o * B + t] = (T)pkg[base + o] / (T)255;
}

// ---- GEMM ----------------------------------------------------------------------
// 64x64 output tile per 256-thread block; thread (ty, tx) owns rows ty+16r and
// columns tx+16s (r, s < 4) so stores are coalesced. Each output is a single
// FMA chain over k in ascending order. For ABT, blockIdx.z selects a k chunk
// and the chunk's partial tile goes to workspace[z].
constexpr int kTile = 64;
constexpr int kKT = 16;

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_gemm(int n, int m, int k, const T* __restrict__ A,
                                              const T* __restrict__ Bm, T* __restrict__ C)
{
  __shared__ T As[kKT][kTile + 1];
  __shared__ T Bs[kKT][kTile + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
  int kbeg = 0, kend = k;
  T* out = C;
  if (OP == NTC_GEMM_ABT)
  {
    kbeg = blockIdx.z * NTC_GEMM_KCHUNK;
    kend = min(k, kbeg + NTC_GEMM_KCHUNK);
    out = C + (int64_t)blockIdx.z * n * m;
  }
  T acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[r][s] = (T)0;

  for (int k0 = kbeg; k0 < kend; k0 += kKT)
  {
    // 64x16 elements of each operand, 4 per thread
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
      const int idx = threadIdx.x + 256 * e;
      int ii, kk;
      if (OP == NTC_GEMM_ATB)
      {
        ii = idx & 63;
        kk = idx >> 6;
      }
      else
      {
        kk = idx & 15;
        ii = idx >> 4;
      }
      const int gi = i0 + ii, gk = k0 + kk;
      T av = (T)0;
      if (gi < n && gk < kend)
        av = OP == NTC_GEMM_ATB ? A[(int64_t)gk * n + gi] : A[(int64_t)gi * k + gk];
      As[kk][ii] = av;

      int jj, kb;
      if (OP == NTC_GEMM_ABT)
      {
        kb = idx & 15;
        jj = idx >> 4;
      }
      else
      {
        jj = idx & 63;
        kb = idx >> 6;
      }
      const int gj = j0 + jj, gkb = k0 + kb;
      T bv = (T)0;
      if (gj < m && gkb < kend)
        bv = OP == NTC_GEMM_ABT ? Bm[(int64_t)gj * k + gkb] : Bm[(int64_t)gkb * m + gj];
      Bs[kb][jj] = bv;
    }
    __syncthreads();
    const int kn = min(kKT, kend - k0);
    for (int kk = 0; kk < kn; ++kk)
    {
      T a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int s = 0; s < 4; ++s) b[s] = Bs[kk][tx + 16 * s];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[r][s] = fma(a[r], b[s], acc[r][s]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
  {
    const int gi = i0 + ty + 16 * r;
    if (gi >= n) continue;
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
      const int gj = j0 + tx + 16 * s;
      if (gj < m) out[(int64_t)gi * m + gj] = acc[r][s];
    }
  }
}

// Sum the k-chunk partials of an ABT product in chunk order.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sum_chunks(int64_t nm, int chunks,
                                                         const T* __restrict__ part,
                                                         T* __restrict__ C)
{
  const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= nm) return;
  T total = (T)0;
  for (int z = 0; z < chunks; ++z) total = total + part[(int64_t)z * nm + e];
  C[e] = total;
}

// ---- activations -------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_act_fwd(int act, int64_t n, const T* __restrict__ z,
                                                      T* __restrict__ a)
{
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    a[i] = act_fwd<T>(act, z[i]);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_act_bwd(int act, int64_t n, const T* __restrict__ z,
                                                      const T* __restrict__ g,
                                                      T* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    out[i] = act_bwd<T>(act, z[i], g[i]);
}

// ---- loss ------------------------------------------------------------------------
constexpr int kLossBlocks = 1024;

template <typename T>
__global__ __launch_bounds__(kThreads) void k_loss(int64_t n, T scale, const T* __restrict__ yt,
                                                   const T* __restrict__ yp, T* __restrict__ dy,
                                                   double* __restrict__ partial)
{
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
  {
    const T r = yt[i] - yp[i];
    if (dy) dy[i] = scale * r;
    s += (double)r * (double)r;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void k_sum_partials(int n, const double* __restrict__ partial,
                                                           double* __restrict__ out)
{
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Per-channel SSE for eval: each block sums O channels over a strided range
// of texels in FP64, then one block adds the block partials in order.
__global__ __launch_bounds__(kThreads) void k_channel_sse(int O, int64_t B, const float* __restrict__ yt,
                                                          const float* __restrict__ yp,
                                                          double* __restrict__ partial)
{
  __shared__ double red[kThreads];
  for (int o = 0; o < O; ++o)
  {
    double s = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < B;
         t += (int64_t)gridDim.x * kThreads)
    {
      const float r = yt[(int64_t)o * B + t] - yp[(int64_t)o * B + t];
      s += (double)r * (double)r;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1)
    {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * 16 + o] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_sum_channel_partials(int nblocks, const double* __restrict__ partial,
                                                              double* __restrict__ sse)
{
  sum_channel_partials(nblocks, partial, sse);
}

// ---- scatter ------------------------------------------------------------------------
// Adjoint of the grid part of the gather: one thread per texel, one atomic add
// per (corner, channel).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter(ntc_feature_desc fd, ntc_tiles tl,
                                                      const T* __restrict__ dx, T* __restrict__ dg0,
                                                      T* __restrict__ dg1, int64_t B)
{
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= B) return;
  const TexelPos p = texel_of(t, tl.origins, tl.tile_dim, tl.mip_dim);
  const Footprint<T> f0 = footprint<T>(p.x, p.y, tl.mip_dim, fd.g0_dim);
  const Footprint<T> f1 = footprint<T>(p.x, p.y, tl.mip_dim, fd.g1_dim);
  const int c0 = fd.g0_channels, c1 = fd.g1_channels;
  const int base1 = fd.g0_mode == NTC_G0_CORNERS ? 4 * c0 : c0;
  for (int c = 0; c < 4; ++c)
  {
    for (int ch = 0; ch < c0; ++ch)
    {
      T v;
      if (fd.g0_mode == NTC_G0_CORNERS)
        v = dx[(int64_t)(c * c0 + ch) * B + t];
      else
        v = f0.w[c] * dx[(int64_t)ch * B + t];
      atomicAdd(&dg0[(int64_t)f0.cell[c] * c0 + ch], v);
    }
    for (int ch = 0; ch < c1; ++ch)
    {
      const T v = f1.w[c] * dx[(int64_t)(base1 + ch) * B + t];
      atomicAdd(&dg1[(int64_t)f1.cell[c] * c1 + ch], v);
    }
  }
}

// ---- host wrappers -------------------------------------------------------------------
template <typename T>
int draw_features(const ntc_feature_desc* fd, const ntc_tiles* tl, const ntc_noise* nz,
                  const T* g0, const T* g1, T* x, hipStream_t s)
{
  NTC_RETURN_IF(!valid_feature_desc(fd) || !valid_tiles(tl) || !g0 || !g1 || !x, NTC_E_INVALID);
  ntc_noise none = {};
  const int64_t B = batch_of(*tl);
  hipLaunchKernelGGL(k_draw_features<T>, dim3(blocks_for(B, kThreads)), dim3(kThreads), 0, s, *fd,
                     *tl, nz ? *nz : none, g0, g1, x, B);
  return NTC_LAUNCH_STATUS();
}

template <typename T>
int draw_targets(const ntc_tiles* tl, int O, const uint8_t* pkg, T* y, hipStream_t s)
{
  NTC_RETURN_IF(!valid_tiles(tl) || O < 1 || O > 16 || !pkg || !y, NTC_E_INVALID);
  const int64_t B = batch_of(*tl);
  hipLaunchKernelGGL(k_draw_targets<T>, dim3(blocks_for(B, kThreads)), dim3(kThreads), 0, s, *tl, O,
                     pkg, y, B);
  return NTC_LAUNCH_STATUS();
}

size_t gemm_ws(int op, int n, int m, int k, int elem)
{
  if (op != NTC_GEMM_ABT) return 0;
  const int64_t chunks = (k + NTC_GEMM_KCHUNK - 1) / NTC_GEMM_KCHUNK;
  return (size_t)chunks * n * m * elem;
}

template <typename T>
int gemm(int op, int n, int m, int k, const T* a, const T* b, T* c, void* ws, size_t ws_bytes,
         hipStream_t s)
{
  NTC_RETURN_IF(n < 1 || m < 1 || k < 1 || !a || !b || !c, NTC_E_INVALID);
  const dim3 grid((m + kTile - 1) / kTile, (n + kTile - 1) / kTile, 1);
  switch (op)
  {
  case NTC_GEMM_AB:
    hipLaunchKernelGGL((k_gemm<T, NTC_GEMM_AB>), grid, dim3(256), 0, s, n, m, k, a, b, c);
    break;
  case NTC_GEMM_ATB:
    hipLaunchKernelGGL((k_gemm<T, NTC_GEMM_ATB>), grid, dim3(256), 0, s, n, m, k, a, b, c);
    break;
  case NTC_GEMM_ABT:
  {
    const size_t need = gemm_ws(op, n, m, k, sizeof(T));
    NTC_RETURN_IF(!ws || ws_bytes < need, NTC_E_WORKSPACE);
    const int chunks = (k + NTC_GEMM_KCHUNK - 1) / NTC_GEMM_KCHUNK;
    const dim3 g3(grid.x, grid.y, chunks);
    hipLaunchKernelGGL((k_gemm<T, NTC_GEMM_ABT>), g3, dim3(256), 0, s, n, m, k, a, b, (T*)ws);
    const int64_t nm = (int64_t)n * m;
    hipLaunchKernelGGL(k_sum_chunks<T>, dim3(blocks_for(nm, kThreads)), dim3(kThreads), 0, s, nm,
                       chunks, (const T*)ws, c);
    break;
  }
  default:
    return NTC_E_INVALID;
  }
  return NTC_LAUNCH_STATUS();
}

unsigned stream_blocks(int64_t n)
{
  const int64_t b = (n + kThreads - 1) / kThreads;
  return (unsigned)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

template <typename T>
int act_forward(int act, int64_t n, const T* z, T* a, hipStream_t s)
{
  NTC_RETURN_IF(n < 1 || !z || !a || (act != NTC_ACT_HARDGELU && act != NTC_ACT_RELU), NTC_E_INVALID);
  hipLaunchKernelGGL(k_act_fwd<T>, dim3(stream_blocks(n)), dim3(kThreads), 0, s, act, n, z, a);
  return NTC_LAUNCH_STATUS();
}

template <typename T>
int act_backward(int act, int64_t n, const T* z, const T* g, T* out, hipStream_t s)
{
  NTC_RETURN_IF(n < 1 || !z || !g || !out || (act != NTC_ACT_HARDGELU && act != NTC_ACT_RELU),
                NTC_E_INVALID);
  hipLaunchKernelGGL(k_act_bwd<T>, dim3(stream_blocks(n)), dim3(kThreads), 0, s, act, n, z, g, out);
  return NTC_LAUNCH_STATUS();
}

template <typename T>
int loss(int64_t n, T scale, const T* yt, const T* yp, T* dy, double* sse, void* ws, hipStream_t s)
{
  NTC_RETURN_IF(n < 1 || !yt || !yp || !sse || !ws, NTC_E_INVALID);
  double* partial = (double*)ws;
  hipLaunchKernelGGL(k_loss<T>, dim3(kLossBlocks), dim3(kThreads), 0, s, n, scale, yt, yp, dy, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, s, kLossBlocks, (const double*)partial,
                     sse);
  return NTC_LAUNCH_STATUS();
}

template <typename T>
int scatter(const ntc_feature_desc* fd, const ntc_tiles* tl, const T* dx, T* dg0, T* dg1,
            hipStream_t s)
{
  NTC_RETURN_IF(!valid_feature_desc(fd) || !valid_tiles(tl) || !dx || !dg0 || !dg1, NTC_E_INVALID);
  const int64_t B = batch_of(*tl);
  hipLaunchKernelGGL(k_scatter<T>, dim3(blocks_for(B, kThreads)), dim3(kThreads), 0, s, *fd, *tl, dx,
                     dg0, dg1, B);
  return NTC_LAUNCH_STATUS();
}
} // namespace

extern "C" {
int ntc_draw_features_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles, const ntc_noise* noise,
                          const float* g0, const float* g1, float* x, hipStream_t stream)
{
  return draw_features<float>(fd, tiles, noise, g0, g1, x, stream);
}

int ntc_draw_features_f64(const ntc_feature_desc* fd, const ntc_tiles* tiles, const ntc_noise* noise,
                          const double* g0, const double* g1, double* x, hipStream_t stream)
{
  return draw_features<double>(fd, tiles, noise, g0, g1, x, stream);
}

int ntc_draw_targets_f32(const ntc_tiles* tiles, int32_t out_dim, const uint8_t* package, float* y,
                         hipStream_t stream)
{
  return draw_targets<float>(tiles, out_dim, package, y, stream);
}

int ntc_draw_targets_f64(const ntc_tiles* tiles, int32_t out_dim, const uint8_t* package, double* y,
                         hipStream_t stream)
{
  return draw_targets<double>(tiles, out_dim, package, y, stream);
}

size_t ntc_gemm_workspace_bytes(int32_t op, int32_t n, int32_t m, int32_t k, int32_t elem_size)
{
  return gemm_ws(op, n, m, k, elem_size);
}

int ntc_gemm_f32(int32_t op, int32_t n, int32_t m, int32_t k, const float* a, const float* b, float* c,
                 void* workspace, size_t workspace_bytes, hipStream_t stream)
{
  return gemm<float>(op, n, m, k, a, b, c, workspace, workspace_bytes, stream);
}

int ntc_gemm_f64(int32_t op, int32_t n, int32_t m, int32_t k, const double* a, const double* b,
                 double* c, void* workspace, size_t workspace_bytes, hipStream_t stream)
{
  return gemm<double>(op, n, m, k, a, b, c, workspace, workspace_bytes, stream);
}

int ntc_act_forward_f32(int32_t act, int64_t n, const float* z, float* a, hipStream_t stream)
{
  return act_forward<float>(act, n, z, a, stream);
}

int ntc_act_backward_f32(int32_t act, int64_t n, const float* z, const float* g, float* out,
                         hipStream_t stream)
{
  return act_backward<float>(act, n, z, g, out, stream);
}

int ntc_act_forward_f64(int32_t act, int64_t n, const double* z, double* a, hipStream_t stream)
{
  return act_forward<double>(act, n, z, a, stream);
}

int ntc_act_backward_f64(int32_t act, int64_t n, const double* z, const double* g, double* out,
                         hipStream_t stream)
{
  return act_backward<double>(act, n, z, g, out, stream);
}

size_t ntc_loss_workspace_bytes(int64_t n)
{
  (void)n;
  return sizeof(double) * kLossBlocks;
}

int ntc_loss_f32(int64_t n, float scale, const float* y_target, const float* y_pred, float* dy,
                 double* sse, void* workspace, hipStream_t stream)
{
  return loss<float>(n, scale, y_target, y_pred, dy, sse, workspace, stream);
}

int ntc_loss_f64(int64_t n, double scale, const double* y_target, const double* y_pred, double* dy,
                 double* sse, void* workspace, hipStream_t stream)
{
  return loss<double>(n, scale, y_target, y_pred, dy, sse, workspace, stream);
}

size_t ntc_channel_sse_workspace_bytes(void) { return sizeof(double) * 16 * kLossBlocks; }

int ntc_channel_sse_f32(int32_t out_dim, int64_t batch, const float* y_target, const float* y_pred,
                        double* sse, void* workspace, hipStream_t stream)
{
  NTC_RETURN_IF(out_dim < 1 || out_dim > 16 || batch < 1 || !y_target || !y_pred || !sse || !workspace,
                NTC_E_INVALID);
  hipLaunchKernelGGL(k_channel_sse, dim3(kLossBlocks), dim3(kThreads), 0, stream, out_dim, batch, y_target,
                     y_pred, (double*)workspace);
  hipLaunchKernelGGL(k_sum_channel_partials, dim3(out_dim), dim3(256), 0, stream, kLossBlocks,
                     (const double*)workspace, sse);
  return NTC_LAUNCH_STATUS();
}

int ntc_accumulate_grid_gradients_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                                      const float* dx, float* dg0, float* dg1, hipStream_t stream)
{
  return scatter<float>(fd, tiles, dx, dg0, dg1, stream);
}

int ntc_accumulate_grid_gradients_f64(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                                      const double* dx, double* dg0, double* dg1, hipStream_t stream)
{
  return scatter<double>(fd, tiles, dx, dg0, dg1, stream);
}
} // extern "C"
