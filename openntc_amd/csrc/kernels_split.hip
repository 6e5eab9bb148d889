// The training step in three launches (gather, MLP, latent scatter) for the
// H = 128 decoder shapes, and its C ABI. The MLP body is in mlp_body.h; the
// gather and the box gradients are the one-launch body's (fused_body.h), run
// here one 128-texel pass per workgroup with no MLP beside them.
#include "decoder_update.h"
#include "grid_lazy.h"
#include "mlp_body.h"

namespace
{
// Pass geometry and LDS staging limits of the gather and scatter launches:
// the member names the shared gather / box helpers read. Every footprint box
// that fits WMAX x VMAX cells is staged (PF quads per thread cover the
// largest), so a blocked pass gathers from LDS whenever its boxes fit.
template <int FP, int G, int C0, int C1, int G0MODE, int PEC>
struct SplitGeo
{
  static constexpr int PASS = kSplitPass;
  static constexpr int BH = PASS / 16;
  static constexpr int KF = FP / 32;
  static constexpr int G1B = G0MODE == NTC_G0_CORNERS ? 4 * C0 : C0;
  static constexpr int GSLOTS = G1B + C1;
  static constexpr int PE_FIRST = PEC == 3 ? 8 : 4;
  static constexpr int XS = FP + 8;
  static constexpr int WMAX = 18, VMAX = BH + 3;
  static constexpr int SW = WMAX, SV = VMAX;
  static constexpr int Q0 = (C0 + 3) / 4, Q1 = (C1 + 3) / 4;
  static constexpr int CP0 = 4 * Q0, CP1 = 4 * Q1;
  static constexpr int PF = (SW * SV * (Q0 + Q1) + kThreads - 1) / kThreads;
  static constexpr int SBOX0 = SW * SV * CP0;
  static constexpr int SBOXF = SW * SV * (CP0 + CP1);  // staged boxes (floats)
  // dX row stride = 16 or 48 mod 64 floats (the smallest such >= G): the
  // four texel rows of one FP32 MFMA B-operand read fall on distinct LDS banks
  static constexpr int GS = G + (((16 - G) % 32) + 32) % 32;
  static_assert(GSLOTS <= G && G <= FP && G % 16 == 0, "grid slots");
  static_assert((WMAX * VMAX + 16) * WMAX < 65536, "small_div range");
};

// One texel's feature slots rounded to BF16, packed two per word into w
// (slot s in the low or high half of w[s / 2]): corner(grid, c, q) gives the
// noisy latents of channel quad q at corner c (corner order and weights of
// the footprint), w0 / w1 the corner weights. The fma order is the per-texel
// gather's (ntc_kernel.cu:75-214 restated in the fused step's gather), so the
// bits match it. Slots past the features stay as the caller set them.
template <class GC, int C0, int C1, int G0MODE, int PEC, int LODS, int NW, class Corner>
__device__ __forceinline__ void texel_features(const FusedParams& P, int px, int py, const float (&w0)[4],
                                               const float (&w1)[4], Corner&& corner, uint32_t (&w)[NW])
{
  static_assert(C0 % 4 == 0 && C1 % 4 == 0, "channel quads fill whole words");
  auto put4 = [&](int s0, float a, float b, float c, float d) {
    w[s0 / 2] = pk_bf16(a, b);
    w[s0 / 2 + 1] = pk_bf16(c, d);
  };
#pragma unroll
  for (int q = 0; q < GC::Q0; ++q)
  {
    float4 c4[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) c4[c] = corner(0, c, q);
    if (G0MODE == NTC_G0_CORNERS)
    {
#pragma unroll
      for (int c = 0; c < 4; ++c) put4(c * C0 + 4 * q, c4[c].x, c4[c].y, c4[c].z, c4[c].w);
    }
    else
    {
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int c = 0; c < 4; ++c)
      {
        acc[0] = fmaf(w0[c], c4[c].x, acc[0]);
        acc[1] = fmaf(w0[c], c4[c].y, acc[1]);
        acc[2] = fmaf(w0[c], c4[c].z, acc[2]);
        acc[3] = fmaf(w0[c], c4[c].w, acc[3]);
      }
      put4(4 * q, acc[0], acc[1], acc[2], acc[3]);
    }
  }
#pragma unroll
  for (int q = 0; q < GC::Q1; ++q)
  {
    float4 c4[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) c4[c] = corner(1, c, q);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
      acc[0] = fmaf(w1[c], c4[c].x, acc[0]);
      acc[1] = fmaf(w1[c], c4[c].y, acc[1]);
      acc[2] = fmaf(w1[c], c4[c].z, acc[2]);
      acc[3] = fmaf(w1[c], c4[c].w, acc[3]);
    }
    put4(GC::G1B + 4 * q, acc[0], acc[1], acc[2], acc[3]);
  }
  // positional encoding and lod: the remaining slots up to 2 NW
  constexpr int NR = 2 * NW - GC::GSLOTS;
  float r[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) r[k] = 0.0f;
  gather_rest<GC, PEC, LODS>(P, px, py, [&](int slot, float f) {
    if (slot < 2 * NW) r[slot - GC::GSLOTS] = f;
  });
#pragma unroll
  for (int k = 0; k < NR; k += 2) w[(GC::GSLOTS + k) / 2] = pk_bf16(r[k], r[k + 1]);
}

// Features of one 128-texel pass per workgroup (one texel per thread) into
// rows of FXS BF16 in HBM (F slots rounded up to 8, the padding zero): each
// thread writes its texel's row as 16-byte stores, a wave 64 consecutive
// rows. A blocked pass whose footprint boxes fit stages the box cells with
// their Philox noise in LDS (one draw per cell and channel quad, shared by
// the texels around the cell); other passes gather each texel from HBM.
constexpr int kGatherThreads = 128;

template <int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS>
__global__ __launch_bounds__(kGatherThreads) void k_split_gather(FusedParams P, uint16_t* __restrict__ xg, int fxs)
{
  using GC = SplitGeo<FP, G, C0, C1, G0MODE, PEC>;
  constexpr int FXS = (GC::GSLOTS + 4 * PEC + LODS + 7) & ~7;  // == fxs
  __shared__ float4 sbox4[GC::SBOXF / 4];
  float* sbox = reinterpret_cast<float*>(sbox4);
  __shared__ AxisTab atab[2];
  __shared__ uint32_t orig_s[128];
  const int tid = threadIdx.x;
  const int64_t pass = blockIdx.x;
  const int T = P.tl.tile_dim, M = P.tl.mip_dim;
  const int D0 = P.fd.g0_dim, D1 = P.fd.g1_dim;
  for (int i = tid; i < 2 * P.tl.n_tiles; i += kGatherThreads) orig_s[i] = (uint32_t)P.tl.origins[i] % (uint32_t)M;
  __syncthreads();
  const PassGeo pg = pass_geo<GC>(pass, orig_s, T, M, D0, D1);
  static_assert(GC::GSLOTS % 2 == 0 && FXS % 8 == 0, "packed rows");
  uint32_t v[FXS / 2];
  const int t = tid;
  if (pg.staged)
  {
    // box cells' channel quads: loads for two items per thread issued
    // together, then their noise and LDS stores
    const int n0 = pg.b0.w * pg.b0.h * GC::Q0;
    const int n = n0 + pg.b1.w * pg.b1.h * GC::Q1;
    for (int e0 = tid; e0 < n; e0 += 2 * kGatherThreads)
    {
      float4 val[2];
      int dst[2], cell[2], gq[2];
#pragma unroll
      for (int k = 0; k < 2; ++k)
      {
        const int e = e0 + k * kGatherThreads;
        const bool live = e < n;
        const bool g1 = live && e >= n0;
        const int ee = !live ? 0 : (g1 ? e - n0 : e);
        const int Q = g1 ? GC::Q1 : GC::Q0, C = g1 ? C1 : C0, CP = g1 ? GC::CP1 : GC::CP0;
        const Box& bx = g1 ? pg.b1 : pg.b0;
        const int D = g1 ? D1 : D0;
        const int q = ee % Q, cb = ee / Q;
        const int vv = small_div(cb, bx.wr), u = cb - vv * bx.w;
        const int cu = bx.u0 + u, cv = bx.v0 + vv;
        const int cx = cu >= D ? cu - D : cu, cy = cv >= D ? cv - D : cv;
        cell[k] = cy * D + cx;
        const float* src = (g1 ? P.g1 : P.g0) + (int64_t)cell[k] * C + 4 * q;
        if (C % 4 == 0)
          val[k] = *reinterpret_cast<const float4*>(src);
        else
          val[k] = make_float4(src[0], 4 * q + 1 < C ? src[1] : 0.0f, 4 * q + 2 < C ? src[2] : 0.0f,
                               4 * q + 3 < C ? src[3] : 0.0f);
        dst[k] = live ? (g1 ? GC::SBOX0 : 0) + cb * CP + 4 * q : -1;
        gq[k] = (int)g1 | (q << 1);
      }
#pragma unroll
      for (int k = 0; k < 2; ++k)
      {
        if (dst[k] < 0) continue;
        const int grid = gq[k] & 1, q = gq[k] >> 1;
        float4 x = val[k];
        if (P.nz.enabled)
        {
          const uint4 w = noise_words(P.nz.seed, P.nz.step, grid, P.nz.level, (uint32_t)cell[k], q);
          const float delta = grid ? P.nz.delta1 : P.nz.delta0;
          x.x = x.x + noise_value(w.x, delta);
          x.y = x.y + noise_value(w.y, delta);
          x.z = x.z + noise_value(w.z, delta);
          x.w = x.w + noise_value(w.w, delta);
        }
        *reinterpret_cast<float4*>(&sbox[dst[k]]) = x;
      }
    }
    if (tid < 64)
      fill_axis_tab<GC::BH>(atab[0], pg.b0, pg.xu0, pg.yu0, M, D0, tid);
    else
      fill_axis_tab<GC::BH>(atab[1], pg.b1, pg.xu0, pg.yu0, M, D1, tid - 64);
    __syncthreads();
    const int x = t & 15, y = t >> 4;
    int off[2][4];
    float w[2][4];
#pragma unroll
    for (int g = 0; g < 2; ++g)
    {
      const AxisTab& tb = atab[g];
      const int bw = g ? pg.b1.w : pg.b0.w, CP = g ? GC::CP1 : GC::CP0, base = g ? GC::SBOX0 : 0;
      const int ax0 = tb.cx0[x], ax1 = tb.cx1[x], ay0 = tb.cy0[y], ay1 = tb.cy1[y];
      const float fx = tb.fx[x], fy = tb.fy[y];
      w[g][0] = (1.0f - fx) * (1.0f - fy);
      w[g][1] = fx * (1.0f - fy);
      w[g][2] = (1.0f - fx) * fy;
      w[g][3] = fx * fy;
      off[g][0] = base + (ay0 * bw + ax0) * CP;
      off[g][1] = base + (ay0 * bw + ax1) * CP;
      off[g][2] = base + (ay1 * bw + ax0) * CP;
      off[g][3] = base + (ay1 * bw + ax1) * CP;
    }
    const int xu = pg.xu0 + x, yu = pg.yu0 + y;
    texel_features<GC, C0, C1, G0MODE, PEC, LODS>(
        P, xu >= M ? xu - M : xu, yu >= M ? yu - M : yu, w[0], w[1],
        [&](int g, int c, int q) { return *reinterpret_cast<const float4*>(&sbox[off[g][c] + 4 * q]); }, v);
  }
  else
  {
    const PassTexel pt = pass_texel(pg.pm, pass, t, GC::PASS, orig_s, T, M, P.B);
#pragma unroll
    for (int s = 0; s < FXS / 2; ++s) v[s] = 0u;
    if (pt.valid)
    {
      const Footprint<float> f0 = footprint<float>(pt.x, pt.y, M, D0);
      const Footprint<float> f1 = footprint<float>(pt.x, pt.y, M, D1);
      texel_features<GC, C0, C1, G0MODE, PEC, LODS>(
          P, pt.x, pt.y, f0.w, f1.w,
          [&](int g, int c, int q) {
            const int C = g ? C1 : C0;
            const int cell = g ? f1.cell[c] : f0.cell[c];
            const float* src = (g ? P.g1 : P.g0) + (int64_t)cell * C + 4 * q;
            float4 x = C % 4 == 0 ? *reinterpret_cast<const float4*>(src)
                                  : make_float4(src[0], 4 * q + 1 < C ? src[1] : 0.0f, 4 * q + 2 < C ? src[2] : 0.0f,
                                                4 * q + 3 < C ? src[3] : 0.0f);
            if (P.nz.enabled)
            {
              const uint4 w = noise_words(P.nz.seed, P.nz.step, g, P.nz.level, (uint32_t)cell, q);
              const float delta = g ? P.nz.delta1 : P.nz.delta0;
              x.x = x.x + noise_value(w.x, delta);
              x.y = x.y + noise_value(w.y, delta);
              x.z = x.z + noise_value(w.z, delta);
              x.w = x.w + noise_value(w.w, delta);
            }
            return x;
          },
          v);
    }
  }
  uint4* dst = reinterpret_cast<uint4*>(xg + (pass * GC::PASS + t) * FXS);
#pragma unroll
  for (int c = 0; c < FXS / 8; ++c) dst[c] = make_uint4(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
  (void)fxs;
}

// Latent gradients of one 128-texel pass per workgroup from its dX rows:
// box gradients (one FP32 MFMA product per 16 cells x 16 channels, one atomic
// per (cell, channel)) for grids whose pass footprint fits a box, else one
// atomic per texel, slot and corner, with the gather's corner weights.
template <int FP, int G, int C0, int C1, int G0MODE, int PEC>
__device__ __forceinline__ void split_scatter_pass(const FusedParams& P, const float* __restrict__ dxg,
                                                   const int64_t pass)
{
  using GC = SplitGeo<FP, G, C0, C1, G0MODE, PEC>;
  __shared__ __attribute__((aligned(16))) float dxs[GC::PASS * GC::GS];
  __shared__ AxisTab atab[2];
  __shared__ uint32_t orig_s[128];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int T = P.tl.tile_dim, M = P.tl.mip_dim;
  const int D0 = P.fd.g0_dim, D1 = P.fd.g1_dim;
  if (tid < 2 * P.tl.n_tiles) orig_s[tid] = (uint32_t)P.tl.origins[tid] % (uint32_t)M;
  {
    const float4* src = reinterpret_cast<const float4*>(dxg + pass * GC::PASS * G);
    constexpr int QR = G / 4;
    for (int c = tid; c < GC::PASS * QR; c += kThreads)
    {
      const int t = c / QR, q = c - t * QR;
      *reinterpret_cast<float4*>(&dxs[t * GC::GS + 4 * q]) = src[c];
    }
  }
  __syncthreads();
  const PassGeo pg = pass_geo<GC>(pass, orig_s, T, M, D0, D1);
  if (tid < 128)
  {
    if (pg.b0.fits) fill_axis_tab<GC::BH>(atab[0], pg.b0, pg.xu0, pg.yu0, M, D0, tid);
  }
  else if (pg.b1.fits)
    fill_axis_tab<GC::BH>(atab[1], pg.b1, pg.xu0, pg.yu0, M, D1, tid - 128);
  __syncthreads();
  if (pg.b0.fits)
    box_gradients<C0, G0MODE == NTC_G0_CORNERS, GC::GS, GC::PASS>(atab[0], pg.b0, dxs, 0, P.dg0, D0, wave, lane);
  if (pg.b1.fits)
  {
    // g1's tiles start on the wave after g0's last one
    const int t0 = pg.b0.fits ? ((pg.b0.w * pg.b0.h + 15) / 16) * ((C0 + 15) / 16) : 0;
    box_gradients<C1, false, GC::GS, GC::PASS>(atab[1], pg.b1, dxs, GC::G1B, P.dg1, D1, (wave + 4 - (t0 & 3)) & 3,
                                               lane);
  }
  if (pg.b0.fits && pg.b1.fits) return;
  for (int item = tid; item < GC::PASS * GC::GSLOTS; item += kThreads)
  {
    const int t = item / GC::GSLOTS, slot = item - t * GC::GSLOTS;
    const bool in_g1 = slot >= GC::G1B;
    if (in_g1 ? pg.b1.fits : pg.b0.fits) continue;
    const PassTexel pt = pass_texel(pg.pm, pass, t, GC::PASS, orig_s, T, M, P.B);
    if (!pt.valid) continue;
    const float dx = dxs[t * GC::GS + slot];
    if (!in_g1)
    {
      const Footprint<float> f0 = footprint<float>(pt.x, pt.y, M, D0);
      if (G0MODE == NTC_G0_CORNERS)
      {
        const int c = slot / C0, ch = slot - c * C0;
        const int cell = c == 0 ? f0.cell[0] : (c == 1 ? f0.cell[1] : (c == 2 ? f0.cell[2] : f0.cell[3]));
        atomicAdd(&P.dg0[(int64_t)cell * C0 + ch], dx);
      }
      else
      {
#pragma unroll
        for (int c = 0; c < 4; ++c) atomicAdd(&P.dg0[(int64_t)f0.cell[c] * C0 + slot], f0.w[c] * dx);
      }
    }
    else
    {
      const Footprint<float> f1 = footprint<float>(pt.x, pt.y, M, D1);
      const int ch = slot - GC::G1B;
#pragma unroll
      for (int c = 0; c < 4; ++c) atomicAdd(&P.dg1[(int64_t)f1.cell[c] * C1 + ch], f1.w[c] * dx);
    }
  }
}

template <int FP, int G, int C0, int C1, int G0MODE, int PEC>
__global__ __launch_bounds__(kThreads) void k_split_scatter(FusedParams P, const float* __restrict__ dxg)
{
  split_scatter_pass<FP, G, C0, C1, G0MODE, PEC>(P, dxg, blockIdx.x);
}

// What follows the MLP launch in one launch: the latent scatter, the next
// step's catch-up of its footprint cells and the decoder update. They touch
// disjoint memory (the scatter: dX rows and latent gradients; the catch-up:
// p, m, v and step counters of cells outside this step's footprints; the
// decoder update: the decoder partials, parameters, moments and BF16 tables),
// so their workgroups may run in any order beside each other, as the same
// kernels on two streams would, without the events between the streams
// (each costs the command processor several microseconds between kernels).
struct TailArgs
{
  int64_t n_scatter;  // scatter workgroups (one pass each)
  int64_t n_catch;    // catch-up workgroups: catch_rb per rectangle slot
  int32_t catch_rb;
  int32_t n_dec_wg;   // decoder-update workgroups (one parameter per thread)
  LazyArgs L;
  int32_t n_dec, n_adam_w01, H, O, blocks, F, gslots;
  const float* part;
  DecAdam c;
  float* grad;
  float* m;
  float* v;
  float* p;
  const double* sse_part;
  double* sse_out;
  TableLayout tly;
  uint16_t* tab;
};

// Workgroups [0, n_dec_wg) update the decoder; the others are dealt to
// scatter passes and catch-up blocks in proportion (workgroup r of them is a
// scatter pass when floor((r + 1) n_scatter / n) steps past floor(r n_scatter
// / n), n = n_scatter + n_catch: every pass once, and the catch-up index is
// r minus the passes before it), so both kinds are in flight throughout.
template <int FP, int G, int C0, int C1, int G0MODE, int PEC>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_split_tail(FusedParams P, const float* __restrict__ dxg, TailArgs T)
{
  const int64_t b = blockIdx.x;
  if (b < T.n_dec_wg)
  {
    decoder_update_elem((int)b * kThreads + (int)threadIdx.x, T.n_dec, T.n_adam_w01, T.H, T.O, T.blocks, T.part,
                        T.c, T.grad, T.m, T.v, T.p, T.sse_part, T.sse_out, T.tly, T.F, T.gslots, T.tab);
    return;
  }
  const int64_t r = b - T.n_dec_wg, n = T.n_scatter + T.n_catch;
  const int64_t s0 = r * T.n_scatter / n, s1 = (r + 1) * T.n_scatter / n;
  if (s1 > s0)
    split_scatter_pass<FP, G, C0, C1, G0MODE, PEC>(P, dxg, s0);
  else
  {
    const int64_t c = r - s0;
    lazy_catchup_rects_body(T.L, (int)(c % T.catch_rb), T.catch_rb, (int)(c / T.catch_rb));
  }
}

// MLP launch: 8 waves (two per SIMD, at most 256 registers each), one
// 16-texel tile per wave, fragments read two ahead of their MFMA (the other
// wave of the SIMD hides the rest of the latency).
constexpr int kMlpWaves = 8, kMlpRing = 2;

template <int H, int FP, int G, int ACT, int FEAT>
__global__ __launch_bounds__(64 * kMlpWaves, 1) void k_mlp_train(SplitParams P)
{
  mlp_body<H, FP, G, ACT, kMlpWaves, kMlpRing, FEAT>(P, blockIdx.x, gridDim.x);
}

typedef void (*GatherKernel)(FusedParams, uint16_t*, int);
typedef void (*ScatterKernel)(FusedParams, const float*);
typedef void (*TailKernel)(FusedParams, const float*, TailArgs);
typedef void (*MlpKernel)(SplitParams);

struct SplitEntry
{
  Shape s;
  GatherKernel gather;
  MlpKernel mlp;
  ScatterKernel scatter;
  TailKernel tail;
};

#define NTC_SPLIT(H, FP, G, C0, C1, MODE, PEC, LODS, ACT)                                                  \
  SplitEntry{shape_v<H, FP, G, C0, C1, MODE, PEC, LODS, ACT>(), k_split_gather<FP, G, C0, C1, MODE, PEC, LODS>, \
             k_mlp_train<H, FP, G, ACT, (MODE == NTC_G0_CORNERS ? 4 * C0 : C0) + C1 + 4 * PEC + LODS>,     \
             k_split_scatter<FP, G, C0, C1, MODE, PEC>, k_split_tail<FP, G, C0, C1, MODE, PEC>},

// Layouts with the three-launch form: the H = 128 decoders with ReLU.
const SplitEntry kSplitEntries[] = {NTC_SPLIT(128, 64, 32, 16, 16, NTC_G0_BILINEAR, 2, 0, NTC_ACT_RELU)};

const SplitEntry* find_split(const ntc_feature_desc& fd, int H, int act)
{
  Shape s;
  shape_of(fd, H, act, s);
  for (const SplitEntry& e : kSplitEntries)
    if (same(e.s, s)) return &e;
  return nullptr;
}

int64_t split_passes(int64_t batch) { return (batch + kSplitPass - 1) / kSplitPass; }

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t split_x_bytes(const ntc_feature_desc& fd, int64_t batch)
{
  return align256((size_t)split_passes(batch) * kSplitPass * split_fxs(fd.feature_dim) * sizeof(uint16_t));
}
} // namespace

extern "C" {
int ntc_split_supported(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t activation)
{
  if (!valid_feature_desc(fd)) return 0;
  return find_split(*fd, hidden_dim, activation) != nullptr;
}

int ntc_split_blocks_for(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t activation, int64_t batch,
                         int32_t max_blocks)
{
  if (!valid_feature_desc(fd) || !find_split(*fd, hidden_dim, activation) || batch < 1 || max_blocks < 1) return 0;
  const int64_t npass = split_passes(batch);
  return (int)(npass < max_blocks ? npass : max_blocks);
}

size_t ntc_split_workspace_bytes(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t activation, int64_t batch)
{
  if (!valid_feature_desc(fd) || batch < 1) return 0;
  const SplitEntry* e = find_split(*fd, hidden_dim, activation);
  if (!e) return 0;
  return split_x_bytes(*fd, batch) + align256((size_t)split_passes(batch) * kSplitPass * e->s.G * sizeof(float));
}

// Checks the arguments of a split stage and fills the launch parameters;
// NTC_OK or the error code.
static int split_prepare(const ntc_fused_args* a, void* workspace, size_t workspace_bytes, const SplitEntry*& e,
                         FusedParams& P, SplitParams& S, int64_t& npass)
{
  NTC_RETURN_IF(!a || !valid_feature_desc(&a->feat) || !valid_tiles(&a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!pow2_tiles(a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!a->g0 || !a->g1 || !a->weights || !a->targets || !a->dw_partials || !a->sse_partials,
                NTC_E_INVALID);
  NTC_RETURN_IF(a->scatter && (!a->dg0 || !a->dg1), NTC_E_INVALID);
  NTC_RETURN_IF(a->out_dim < 1 || a->out_dim > 16 || a->blocks < 1, NTC_E_INVALID);
  e = find_split(a->feat, a->hidden_dim, a->activation);
  NTC_RETURN_IF(!e, NTC_E_INVALID);
  const int64_t B = batch_of(a->tiles);
  npass = split_passes(B);
  NTC_RETURN_IF(npass > 0x7FFFFFFF, NTC_E_INVALID);
  const int blocks = (int)(npass < a->blocks ? npass : a->blocks);
  NTC_RETURN_IF(blocks != a->blocks, NTC_E_INVALID);  // every partial slot must be written
  NTC_RETURN_IF(!workspace || workspace_bytes < ntc_split_workspace_bytes(&a->feat, a->hidden_dim, a->activation, B),
                NTC_E_WORKSPACE);
  NTC_RETURN_IF(((uintptr_t)workspace & 255) != 0, NTC_E_INVALID);
  uint16_t* xg = static_cast<uint16_t*>(workspace);
  float* dxg = reinterpret_cast<float*>(static_cast<char*>(workspace) + split_x_bytes(a->feat, B));
  P = {};
  P.fd = a->feat;
  P.tl = a->tiles;
  P.nz = a->noise;
  P.g0 = a->g0;
  P.g1 = a->g1;
  P.wtab = a->weights;
  P.pkg = a->targets;
  P.O = a->out_dim;
  P.F = a->feat.feature_dim;
  P.scale = a->loss_scale;
  P.B = B;
  P.scatter = a->scatter;
  P.dg0 = a->dg0;
  P.dg1 = a->dg1;
  S = {};
  S.tl = a->tiles;
  S.xg = xg;
  S.dxg = dxg;
  S.wtab = a->weights;
  S.pkg = a->targets;
  S.O = a->out_dim;
  S.F = a->feat.feature_dim;
  S.scale = a->loss_scale;
  S.B = B;
  S.dw_part = a->dw_partials;
  S.sse_part = a->sse_partials;
  S.dec_elems = a->hidden_dim * a->feat.feature_dim + a->hidden_dim * a->hidden_dim + 16 * a->hidden_dim;
  return NTC_OK;
}

int ntc_split_gather(const ntc_fused_args* a, void* workspace, size_t workspace_bytes, hipStream_t stream)
{
  const SplitEntry* e;
  FusedParams P;
  SplitParams S;
  int64_t npass;
  const int r = split_prepare(a, workspace, workspace_bytes, e, P, S, npass);
  if (r != NTC_OK) return r;
  hipLaunchKernelGGL(e->gather, dim3((unsigned)npass), dim3(kGatherThreads), 0, stream, P,
                     static_cast<uint16_t*>(workspace), split_fxs(a->feat.feature_dim));
  return NTC_LAUNCH_STATUS();
}

int ntc_split_mlp(const ntc_fused_args* a, void* workspace, size_t workspace_bytes, hipStream_t stream)
{
  const SplitEntry* e;
  FusedParams P;
  SplitParams S;
  int64_t npass;
  const int r = split_prepare(a, workspace, workspace_bytes, e, P, S, npass);
  if (r != NTC_OK) return r;
  hipLaunchKernelGGL(e->mlp, dim3(a->blocks), dim3(64 * kMlpWaves), 0, stream, S);
  return NTC_LAUNCH_STATUS();
}

int ntc_split_scatter(const ntc_fused_args* a, void* workspace, size_t workspace_bytes, hipStream_t stream)
{
  const SplitEntry* e;
  FusedParams P;
  SplitParams S;
  int64_t npass;
  const int r = split_prepare(a, workspace, workspace_bytes, e, P, S, npass);
  if (r != NTC_OK) return r;
  if (a->scatter) hipLaunchKernelGGL(e->scatter, dim3((unsigned)npass), dim3(kThreads), 0, stream, P, S.dxg);
  return NTC_LAUNCH_STATUS();
}

int ntc_split_tail(const ntc_fused_args* a, void* workspace, size_t workspace_bytes, const ntc_split_tail_args* t,
                   hipStream_t stream)
{
  const SplitEntry* e;
  FusedParams P;
  SplitParams S;
  int64_t npass;
  int r = split_prepare(a, workspace, workspace_bytes, e, P, S, npass);
  if (r != NTC_OK) return r;
  NTC_RETURN_IF(!t, NTC_E_INVALID);
  TailArgs T = {};
  T.n_scatter = a->scatter ? npass : 0;
  if (t->catch_tiles && t->catch_upto > 0)
  {
    // the catch-up touches no gradient: any non-null pointer passes the checks
    r = lazy_args_of(t->catch_feat, t->catch_tiles, t->catch_g0, t->catch_g0, t->catch_g1, t->catch_g1,
                     t->catch_adam, t->catch_steps, t->catch_upto, t->catch_last0, t->catch_last1, T.L);
    if (r != NTC_OK) return r;
    if (t->catch_skip)
    {
      NTC_RETURN_IF(!valid_tiles(t->catch_skip) || t->catch_skip->n_tiles > 16, NTC_E_INVALID);
      T.L.skip = *t->catch_skip;
      T.L.n_skip = t->catch_skip->n_tiles;
    }
    T.catch_rb = rect_blocks(t->catch_tiles, T.L.A, 256);
    T.n_catch = (int64_t)T.catch_rb * 2 * t->catch_tiles->n_tiles;
  }
  if (t->dec_p)
  {
    const ntc_feature_desc* fd = t->dec_feat;
    const int H = a->hidden_dim, O = a->out_dim;
    NTC_RETURN_IF(!fd || !valid_feature_desc(fd) || fd->feature_dim != a->feat.feature_dim, NTC_E_INVALID);
    NTC_RETURN_IF(!t->dec_grad || !t->dec_m || !t->dec_v || !t->dec_table, NTC_E_INVALID);
    Shape s;
    shape_of(*fd, H, 0, s);
    const int F = fd->feature_dim;
    T.n_dec = H * F + H * H + 16 * H;
    T.n_adam_w01 = H * F + H * H;
    T.H = H;
    T.O = O;
    T.blocks = a->blocks;
    T.F = F;
    T.gslots = g1_base(*fd) + fd->g1_channels;
    T.part = a->dw_partials;
    T.c = DecAdam{t->dec_lr, t->dec_beta1, t->dec_beta2, 1.0f - t->dec_beta1, 1.0f - t->dec_beta2, t->dec_bias1,
                  t->dec_bias2};
    T.grad = t->dec_grad;
    T.m = t->dec_m;
    T.v = t->dec_v;
    T.p = t->dec_p;
    T.sse_part = a->sse_partials;
    T.sse_out = t->sse_out;
    T.tly = TableLayout{s.H, s.FP, s.G};
    T.tab = t->dec_table;
    T.n_dec_wg = (T.n_dec + kThreads - 1) / kThreads;
  }
  const int64_t grid = T.n_dec_wg + T.n_scatter + T.n_catch;
  NTC_RETURN_IF(grid > 0x7FFFFFFF, NTC_E_INVALID);
  if (grid == 0) return NTC_OK;
  hipLaunchKernelGGL(e->tail, dim3((unsigned)grid), dim3(kThreads), 0, stream, P, S.dxg, T);
  return NTC_LAUNCH_STATUS();
}

int ntc_split_train_step(const ntc_fused_args* a, void* workspace, size_t workspace_bytes, hipStream_t stream)
{
  int r = ntc_split_gather(a, workspace, workspace_bytes, stream);
  if (r == NTC_OK) r = ntc_split_mlp(a, workspace, workspace_bytes, stream);
  if (r == NTC_OK) r = ntc_split_scatter(a, workspace, workspace_bytes, stream);
  return r;
}
} // extern "C"
