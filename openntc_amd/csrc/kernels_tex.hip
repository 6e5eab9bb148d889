// Mip-target kernels: channel selection, separable 2:1 Lanczos-3 downsample
// and interleaving into the per-mip target package.
//
//   k_prepare_tex    ntc_kernel.cu:533-567 (prepare_tex)
//   k_lanczos_x/_y   ntc_kernel.cu:569-663 (filter_lanczos, two passes)
//   k_package_tex    ntc_kernel.cu:665-686 (package_tex)
//
// The 12 filter weights L((i+0.5)/2), i in [-6, 5], L(x) = sinc(pi x)
// sinc(pi x/3), are evaluated once on the host in double precision and rounded
// to float; each output texel accumulates with FMA in tap order, skipping taps
// outside the image and renormalising by the summed weight, then the second
// pass truncates to uint8 like the reference.
#include <math.h>

#include "ntc_device.h"
#include "ntc_host.h"

using namespace ntc;

namespace
{
struct LanczosTaps
{
  float w[12];
};

LanczosTaps make_taps()
{
  LanczosTaps t;
  const double pi = 3.14159265358979323846;
  for (int i = -6; i < 6; ++i)
  {
    const double x = (0.5 + i) * 0.5;
    const double a = pi * x, b = pi * x / 3.0;
    const double sa = fabs(a) < 1e-8 ? 1.0 : sin(a) / a;
    const double sb = fabs(b) < 1e-8 ? 1.0 : sin(b) / b;
    t.w[i + 6] = (float)(sa * sb);
  }
  return t;
}

__global__ __launch_bounds__(256) void k_prepare_tex(int dim, int c_src, int count, int4 cmap,
                                                     const uint8_t* __restrict__ tex,
                                                     uint8_t* __restrict__ out)
{
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= dim || y >= dim) return;
  const int64_t src = ((int64_t)y * dim + x) * c_src;
  const int64_t dst = ((int64_t)y * dim + x) * count;
  const int m[4] = {cmap.x, cmap.y, cmap.z, cmap.w};
  for (int i = 0; i < count; ++i) out[dst + i] = tex[src + m[i]];
}

// pass 1 (x): src uint8 [dim][dim][c] -> tmp float [dim][dim/2][c]
__global__ __launch_bounds__(256) void k_lanczos_x(int dim, int c, LanczosTaps taps,
                                                   const uint8_t* __restrict__ src,
                                                   float* __restrict__ tmp)
{
  const int half = dim / 2;
  const int ox = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (ox >= half || y >= dim) return;
  float tot[4] = {0.f, 0.f, 0.f, 0.f};
  float tw = 0.0f;
  for (int i = -6; i < 6; ++i)
  {
    const int s = 2 * ox + 1 + i;
    if (s < 0 || s >= dim) continue;
    const float w = taps.w[i + 6];
    for (int ch = 0; ch < c; ++ch)
      tot[ch] = fmaf((float)src[((int64_t)y * dim + s) * c + ch] / 255.0f, w, tot[ch]);
    tw = tw + w;
  }
  for (int ch = 0; ch < c; ++ch) tmp[((int64_t)y * half + ox) * c + ch] = tot[ch] / tw;
}

// pass 2 (y): tmp float [dim][dim/2][c] -> dst uint8 [dim/2][dim/2][c], truncating
__global__ __launch_bounds__(256) void k_lanczos_y(int dim, int c, LanczosTaps taps,
                                                   const float* __restrict__ tmp,
                                                   uint8_t* __restrict__ dst)
{
  const int half = dim / 2;
  const int ox = blockIdx.x * 16 + (threadIdx.x & 15);
  const int oy = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (ox >= half || oy >= half) return;
  float tot[4] = {0.f, 0.f, 0.f, 0.f};
  float tw = 0.0f;
  for (int i = -6; i < 6; ++i)
  {
    const int s = 2 * oy + 1 + i;
    if (s < 0 || s >= dim) continue;
    const float w = taps.w[i + 6];
    for (int ch = 0; ch < c; ++ch) tot[ch] = fmaf(tmp[((int64_t)s * half + ox) * c + ch], w, tot[ch]);
    tw = tw + w;
  }
  for (int ch = 0; ch < c; ++ch)
  {
    const float v = tot[ch] / tw;
    const float q = fminf(255.0f, fmaxf(0.0f, v * 255.0f));
    dst[((int64_t)oy * half + ox) * c + ch] = (uint8_t)q;
  }
}

struct PackageArgs
{
  const uint8_t* src[16];
  int src_channels[16];
  int feat_source[16];
  int feat_channel[16];
};

__global__ __launch_bounds__(256) void k_package_tex(int dim, int c, PackageArgs a,
                                                     uint8_t* __restrict__ out)
{
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= dim || y >= dim) return;
  for (int i = 0; i < c; ++i)
  {
    const int sid = a.feat_source[i];
    const int sc = a.src_channels[sid];
    out[((int64_t)y * dim + x) * c + i] = a.src[sid][((int64_t)y * dim + x) * sc + a.feat_channel[i]];
  }
}
} // namespace

extern "C" {
int ntc_prepare_tex(int32_t dim, int32_t src_channels, const int32_t* cmap, const uint8_t* tex,
                    uint8_t* out, hipStream_t stream)
{
  NTC_RETURN_IF(dim < 1 || src_channels < 1 || src_channels > 4 || !cmap || !tex || !out, NTC_E_INVALID);
  int count = 0;
  for (int i = 0; i < 4; ++i)
  {
    if (cmap[i] < 0 || cmap[i] >= src_channels) break;
    ++count;
  }
  NTC_RETURN_IF(count == 0, NTC_E_INVALID);
  const int4 m = make_int4(cmap[0], count > 1 ? cmap[1] : 0, count > 2 ? cmap[2] : 0,
                           count > 3 ? cmap[3] : 0);
  const dim3 grid((dim + 15) / 16, (dim + 15) / 16);
  hipLaunchKernelGGL(k_prepare_tex, grid, dim3(256), 0, stream, dim, src_channels, count, m, tex, out);
  return NTC_LAUNCH_STATUS();
}

int ntc_filter_lanczos(int32_t dim_src, int32_t channels, const uint8_t* src, float* tmp, uint8_t* dst,
                       hipStream_t stream)
{
  NTC_RETURN_IF(dim_src < 2 || (dim_src & 1) || channels < 1 || channels > 4, NTC_E_INVALID);
  NTC_RETURN_IF(!src || !tmp || !dst, NTC_E_INVALID);
  const LanczosTaps taps = make_taps();
  const int half = dim_src / 2;
  hipLaunchKernelGGL(k_lanczos_x, dim3((half + 15) / 16, (dim_src + 15) / 16), dim3(256), 0, stream,
                     dim_src, channels, taps, src, tmp);
  hipLaunchKernelGGL(k_lanczos_y, dim3((half + 15) / 16, (half + 15) / 16), dim3(256), 0, stream, dim_src,
                     channels, taps, (const float*)tmp, dst);
  return NTC_LAUNCH_STATUS();
}

int ntc_package_tex(int32_t dim, int32_t out_channels, int32_t n_sources, const uint8_t* const* sources,
                    const int32_t* source_channels, const int32_t* feat_source,
                    const int32_t* feat_channel, uint8_t* out, hipStream_t stream)
{
  NTC_RETURN_IF(dim < 1 || out_channels < 1 || out_channels > 16 || n_sources < 1 || n_sources > 16,
                NTC_E_INVALID);
  NTC_RETURN_IF(!sources || !source_channels || !feat_source || !feat_channel || !out, NTC_E_INVALID);
  PackageArgs a = {};
  for (int i = 0; i < n_sources; ++i)
  {
    NTC_RETURN_IF(!sources[i] || source_channels[i] < 1 || source_channels[i] > 4, NTC_E_INVALID);
    a.src[i] = sources[i];
    a.src_channels[i] = source_channels[i];
  }
  for (int i = 0; i < out_channels; ++i)
  {
    NTC_RETURN_IF(feat_source[i] < 0 || feat_source[i] >= n_sources, NTC_E_INVALID);
    NTC_RETURN_IF(feat_channel[i] < 0 || feat_channel[i] >= source_channels[feat_source[i]], NTC_E_INVALID);
    a.feat_source[i] = feat_source[i];
    a.feat_channel[i] = feat_channel[i];
  }
  const dim3 grid((dim + 15) / 16, (dim + 15) / 16);
  hipLaunchKernelGGL(k_package_tex, grid, dim3(256), 0, stream, dim, out_channels, a, out);
  return NTC_LAUNCH_STATUS();
}
} // extern "C"
