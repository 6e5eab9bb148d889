// The training step in three launches for the large decoder shapes (H = 128):
//   k_split_gather   features of every texel of the step (staged footprint
//                    boxes with their Philox noise, the fused step's gather)
//                    as BF16 rows in HBM;
//   k_mlp_train      forward + loss + hand-written backward + per-workgroup
//                    decoder gradients on BF16 MFMA, 32 texels per wave; reads
//                    the feature rows, writes dX of the grid slots (FP32);
//   k_split_scatter  latent gradients from the dX rows: per pass footprint box
//                    one FP32 MFMA product per 16 cells x 16 channels and one
//                    atomic per (cell, channel), as the fused step does.
// Same arithmetic and rounding points as the one-launch body of fused_body.h
// (features and activations rounded to BF16, FP32 accumulation, dX in FP32),
// so both forms are held to the same checks. Taking the gather and the
// scatter out of the MLP launch leaves its registers and LDS to the MLP: the
// weight fragments read from LDS feed two MFMAs each (two 16-texel column
// tiles per wave) instead of one.
// Reference stages replaced: libopenntc.cpp:621-709, ntc_kernel.cu:75-373.
#pragma once

#include "fused_body.h"

namespace
{
// Passes of 128 texels (16 x 8 blocks of a tile for tiles of side >= 16);
// the three launches share the pass decomposition, and the feature and dX
// rows are stored in pass order (row pass * 128 + texel of the pass).
constexpr int kSplitPass = 128;

struct SplitParams
{
  ntc_tiles tl;
  const uint16_t* xg;   // feature rows [rows][FXS] BF16, pass order
  float* dxg;           // dX rows [rows][G] FP32, pass order
  const uint16_t* wtab;
  const uint8_t* pkg;
  int O;
  int F;
  float scale;
  int64_t B;
  float* dw_part;
  double* sse_part;
  int dec_elems;
};

// W waves per workgroup, each with NT 16-texel column tiles: W * NT * 16 =
// kSplitPass. The decoder-gradient rounds take 64 texels (four tiles) each;
// every wave accumulates H / W rows of dW1 and dW0 and H / (16 W) column
// tiles of dWout.
template <int H, int FP, int G, int ACT, int W>
struct MlpCfg
{
  static constexpr int NW = W;
  static constexpr int THREADS = 64 * W;
  static constexpr int NT = kSplitPass / (16 * W);
  static constexpr int PASS = kSplitPass;
  static constexpr int ROUNDS = PASS / 64;
  static constexpr int MT = H / 16, KF = FP / 32, KH = H / 32, GT = G / 16;
  static constexpr int RW = H / W, MTW = RW / 16, NTO = H / (16 * W);
  static_assert(NT >= 1 && MTW >= 1 && NTO >= 1 && (W == 4 || NT == 1), "wave shapes");
  static constexpr int WT = H * FP + H * H + 16 * H;
  static constexpr int XS = FP + 8;      // LDS X row stride (elements)
  static constexpr int XR = PASS * XS;
  static constexpr int SH = H + 16;      // staging row stride, as in the one-launch body
  static constexpr int STAGE = 4 * 64 * SH + 64 * 16;
  static constexpr int LDS = WT + XR + STAGE;  // BF16 elements
  static_assert(2 * LDS + 16 * PASS <= 160 * 1024, "LDS budget");
  static_assert(STAGE * 2 >= 8 * 256, "final reduction reuses the staging region");
};

// Feature row stride in HBM: F rounded up to 8 elements (16-byte chunks).
__host__ __device__ constexpr int split_fxs(int F) { return (F + 7) & ~7; }

template <int H, int FP, int G, int ACT, int W, int RING>
__device__ __forceinline__ void mlp_body(const SplitParams& P, const int vblk, const int vgrid)
{
  using CF = MlpCfg<H, FP, G, ACT, W>;
  constexpr int kNT = CF::NT, kPass = CF::PASS, kThreads = CF::THREADS;
  constexpr int MT = CF::MT, KF = CF::KF, KH = CF::KH, GT = CF::GT, MTW = CF::MTW, NTO = CF::NTO;
  constexpr int kRing = RING;
  constexpr TableLayout TLY = {H, FP, G};
  __shared__ __attribute__((aligned(16))) uint16_t smem[CF::LDS];
  uint16_t* wt = smem;
  uint16_t* xnf = smem + CF::WT;
  uint16_t* stA = xnf + CF::XR;
  uint16_t* stB = stA + 64 * CF::SH;
  uint16_t* stC = stB + 64 * CF::SH;
  uint16_t* stD = stC + 64 * CF::SH;
  uint16_t* stY = stD + 64 * CF::SH;
  __shared__ uint32_t orig_s[128];
  __shared__ uint32_t tgt_s[kPass * 4];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int n16 = lane & 15, g = lane >> 4;
  const int FXS = split_fxs(P.F);
  const int xchunks = FXS / 8;  // 16-byte chunks per feature row

  {
    const uint4* src = reinterpret_cast<const uint4*>(P.wtab);
    uint4* dst = reinterpret_cast<uint4*>(wt);
    for (int i = tid; i < CF::WT / 8; i += kThreads) dst[i] = src[i];
    // X row columns past the stored features stay zero for the whole launch
    for (int i = tid; i < kPass * CF::XS; i += kThreads)
      if (i % CF::XS >= FXS) xnf[i] = 0;
  }

  f32x4 dw1[MTW][MT], dw0[MTW][FP / 16], dwo[NTO];
#pragma unroll
  for (int a = 0; a < MTW; ++a)
  {
#pragma unroll
    for (int c = 0; c < MT; ++c) dw1[a][c] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < FP / 16; ++c) dw0[a][c] = f32x4{0, 0, 0, 0};
  }
#pragma unroll
  for (int c = 0; c < NTO; ++c) dwo[c] = f32x4{0, 0, 0, 0};
  double sse = 0.0;

  const int64_t npass = (P.B + kPass - 1) / kPass;
  const int T = P.tl.tile_dim, M = P.tl.mip_dim;
  if (tid < 2 * P.tl.n_tiles) orig_s[tid] = (uint32_t)P.tl.origins[tid] % (uint32_t)M;
  __syncthreads();

  // A pass's feature rows: (kPass * FXS / 8) 16-byte chunks (FXS <= 48),
  // loaded into registers while the previous pass computes its decoder
  // gradients and written into the X rows after them.
  constexpr int kXL = (kPass * 6 + kThreads - 1) / kThreads;
  uint4 xr[kXL];
  auto x_issue = [&](int64_t pass) {
    const uint4* src = reinterpret_cast<const uint4*>(P.xg + pass * kPass * FXS);
#pragma unroll
    for (int k = 0; k < kXL; ++k)
    {
      const int c = tid + k * kThreads;
      xr[k] = src[c < kPass * xchunks ? c : 0];
    }
  };
  auto x_commit = [&]() {
#pragma unroll
    for (int k = 0; k < kXL; ++k)
    {
      const int c = tid + k * kThreads;
      if (c < kPass * xchunks)
      {
        const int t = c / xchunks, q = c - t * xchunks;
        *reinterpret_cast<uint4*>(&xnf[t * CF::XS + 8 * q]) = xr[k];
      }
    }
  };
  {
    uint32_t tg[kNT][4];
    const PassMap pm = pass_map((int64_t)vblk, T, kPass);
    x_issue(vblk);
    targets_issue<CF, kNT>(tg, pm, (int64_t)vblk, P, orig_s, wave, n16, g);
    x_commit();
    targets_commit(tg, tgt_s, wave, n16, g);
  }

  for (int64_t pass = vblk; pass < npass; pass += vgrid)
  {
    __syncthreads();  // this pass's X rows and targets are in LDS
    const PassMap pm = pass_map(pass, T, kPass);
    int tloc[kNT];
    bool tvalid[kNT];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt)
    {
      tloc[nt] = wave * 16 * kNT + nt * 16 + n16;
      tvalid[nt] = pm.blocked || pass * kPass + tloc[nt] < P.B;
    }
    // layer 0
    f32x4 z0[MT][kNT];
    {
      bf16x8 xb[KF][kNT];
#pragma unroll
      for (int j = 0; j < KF; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt) xb[j][nt] = frag_lds(&xnf[tloc[nt] * CF::XS + 32 * j + 8 * g]);
      lds_ring<MT * KF, kRing>(
          [&](int i) {
            const int row = 16 * (i % MT) + n16;
            return &wt[TLY.oW0() + row * FP + swz(row, 32 * (i / MT) + 8 * g, FP)];
          },
          [&](int i, const bf16x8 a) {
            const int r = i % MT, j = i / MT;
            if (j == 0)
#pragma unroll
              for (int nt = 0; nt < kNT; ++nt) z0[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) z0[r][nt] = mfma(a, xb[j][nt], z0[r][nt]);
          });
    }
    bf16x8 a0f[KH][kNT];
#pragma unroll
    for (int j = 0; j < KH; ++j)
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
      {
        float v[8];
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
          v[i] = act_fwd<float>(ACT, z0[2 * j][nt][i]);
          v[4 + i] = act_fwd<float>(ACT, z0[2 * j + 1][nt][i]);
        }
        a0f[j][nt] = frag_pack(v);
      }
    // layer 1
    f32x4 z1[MT][kNT];
    lds_ring<MT * KH, kRing>(
        [&](int i) {
          const int row = 16 * (i % MT) + n16;
          return &wt[TLY.oW1() + row * H + swz(row, 32 * (i / MT) + 8 * g, H)];
        },
        [&](int i, const bf16x8 a) {
          const int r = i % MT, j = i / MT;
          if (j == 0)
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) z1[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
          for (int nt = 0; nt < kNT; ++nt) z1[r][nt] = mfma(a, a0f[j][nt], z1[r][nt]);
        });
    bf16x8 a1f[KH][kNT];
#pragma unroll
    for (int j = 0; j < KH; ++j)
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
      {
        float v[8];
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
          v[i] = act_fwd<float>(ACT, z1[2 * j][nt][i]);
          v[4 + i] = act_fwd<float>(ACT, z1[2 * j + 1][nt][i]);
        }
        a1f[j][nt] = frag_pack(v);
      }
    // output layer + loss
    float dy[kNT][4];
    bf16x8 dyf[kNT];
    {
      f32x4 yh[kNT];
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) yh[nt] = f32x4{0, 0, 0, 0};
      lds_ring<KH, kRing>([&](int j) { return &wt[TLY.oWo() + n16 * H + swz(n16, 32 * j + 8 * g, H)]; },
                          [&](int j, const bf16x8 a) {
#pragma unroll
                            for (int nt = 0; nt < kNT; ++nt) yh[nt] = mfma(a, a1f[j][nt], yh[nt]);
                          });
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
      {
        float v[8];
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
          const int o = 4 * g + i;
          float d = 0.0f;
          if (tvalid[nt] && o < P.O)
          {
            const float yt = (float)((tgt_s[tloc[nt] * 4 + g] >> (8 * i)) & 0xFFu) / 255.0f;
            const float r = yt - yh[nt][i];
            sse += (double)r * (double)r;
            d = P.scale * r;
          }
          dy[nt][i] = d;
          v[i] = d;
          v[4 + i] = 0.0f;
        }
        dyf[nt] = frag_pack(v);
      }
    }
    // backward through the output layer and the activation of layer 1
    bf16x8 dz1f[KH][kNT];
    {
      f32x4 dz1[MT][kNT];
      lds_ring_ld<MT, kRing>(
          [&](int r) {
            const int o = 4 * g + (n16 >> 2), c = kperm_col(16 * r + 4 * (n16 & 3));
            return frag_tr_half(&wt[TLY.oWo() + o * H + swz(o, c, H)]);
          },
          [&](int r, const bf16x8 a) {
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt)
            {
              const f32x4 da = mfma(a, dyf[nt], f32x4{0, 0, 0, 0});
#pragma unroll
              for (int i = 0; i < 4; ++i)
                dz1[r][nt][i] = act_bwd<float>(ACT, pre_act<ACT>(z1, a1f, r, nt, i), da[i]);
            }
          });
#pragma unroll
      for (int j = 0; j < KH; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
        {
          float v[8];
#pragma unroll
          for (int i = 0; i < 4; ++i)
          {
            v[i] = dz1[2 * j][nt][i];
            v[4 + i] = dz1[2 * j + 1][nt][i];
          }
          dz1f[j][nt] = frag_pack(v);
        }
    }
    // backward through layer 1 and the activation of layer 0
    bf16x8 dz0f[KH][kNT];
    {
      f32x4 dz0[MT][kNT];
      lds_ring_ld<MT * KH, kRing>(
          [&](int i) {
            const int j = i / MT, c = kperm_col(16 * (i % MT) + 4 * (n16 & 3));
            const int r0 = tr_row(j, 0, g, n16), r1 = tr_row(j, 1, g, n16);
            return frag_tr(&wt[TLY.oW1() + r0 * H + swz(r0, c, H)], &wt[TLY.oW1() + r1 * H + swz(r1, c, H)]);
          },
          [&](int i, const bf16x8 a) {
            const int r = i % MT, j = i / MT;
            if (j == 0)
#pragma unroll
              for (int nt = 0; nt < kNT; ++nt) dz0[r][nt] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) dz0[r][nt] = mfma(a, dz1f[j][nt], dz0[r][nt]);
            if (j == KH - 1)
#pragma unroll
              for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                  dz0[r][nt][e] = act_bwd<float>(ACT, pre_act<ACT>(z0, a0f, r, nt, e), dz0[r][nt][e]);
          });
#pragma unroll
      for (int j = 0; j < KH; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
        {
          float v[8];
#pragma unroll
          for (int i = 0; i < 4; ++i)
          {
            v[i] = dz0[2 * j][nt][i];
            v[4 + i] = dz0[2 * j + 1][nt][i];
          }
          dz0f[j][nt] = frag_pack(v);
        }
    }
    // dX of the grid slots (FP32) into the pass's dX rows: lane l holds
    // slots 16 r + 4 g .. + 3 of texel l % 16, one 16-byte store
    {
      f32x4 dx[GT][kNT];
      lds_ring_ld<GT * KH, kRing>(
          [&](int q) {
            const int j = q % KH, c = 16 * (q / KH) + 4 * (n16 & 3);
            const int r0 = tr_row(j, 0, g, n16), r1 = tr_row(j, 1, g, n16);
            return frag_tr(&wt[TLY.oW0() + r0 * FP + swz(r0, c, FP)], &wt[TLY.oW0() + r1 * FP + swz(r1, c, FP)]);
          },
          [&](int q, const bf16x8 a) {
            const int r = q / KH, j = q % KH;
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) dx[r][nt] = mfma(a, dz0f[j][nt], j == 0 ? f32x4{0, 0, 0, 0} : dx[r][nt]);
          });
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
      {
        float* drow = P.dxg + (pass * kPass + tloc[nt]) * G;
#pragma unroll
        for (int r = 0; r < GT; ++r) *reinterpret_cast<f32x4*>(drow + 16 * r + 4 * g) = dx[r][nt];
      }
    }

    // next pass's feature rows and targets: the loads land during the rounds
    const int64_t next = pass + vgrid;
    const int64_t pn = next < npass ? next : pass;
    uint32_t tg[kNT][4];
    x_issue(pn);
    targets_issue<CF, kNT>(tg, pass_map(pn, T, kPass), pn, P, orig_s, wave, n16, g);
#pragma unroll
    for (int rnd = 0; rnd < CF::ROUNDS; ++rnd)
    {
      // the round's four tiles: tile rnd of every wave (W = 4), or the one
      // tile of waves 4 rnd .. 4 rnd + 3 (NT = 1)
      const int st = kNT == 1 ? 0 : rnd;
      if (kNT > 1 || (wave >> 2) == rnd)
      {
        const int row = (wave & 3) * 16 + n16;
#pragma unroll
        for (int j = 0; j < KH; ++j)
        {
          const int at = row * CF::SH + 32 * j + 8 * g;
          *reinterpret_cast<bf16x8*>(&stA[at]) = a0f[j][st];
          *reinterpret_cast<bf16x8*>(&stB[at]) = dz1f[j][st];
          *reinterpret_cast<bf16x8*>(&stC[at]) = dz0f[j][st];
          *reinterpret_cast<bf16x8*>(&stD[at]) = a1f[j][st];
        }
        *reinterpret_cast<uint2*>(&stY[row * 16 + 4 * g]) =
            make_uint2(pk_bf16(dy[st][0], dy[st][1]), pk_bf16(dy[st][2], dy[st][3]));
      }
      __syncthreads();
      const int trr0 = 8 * g + (n16 >> 2), tq = 4 * (n16 & 3);
      auto frag_st = [&](const uint16_t* st, int stride, int kc, int c0) {
        const uint16_t* p = st + (32 * kc + trr0) * stride + c0 + tq;
        return frag_tr(p, p + 4 * stride);
      };
      // dW1[p_out][p_in] += dZ1 A0^T; each A0^T fragment read once and used
      // for all of this wave's row tiles
      {
        bf16x8 a[MTW][2];
#pragma unroll
        for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
          for (int kc = 0; kc < 2; ++kc) a[mt][kc] = frag_st(stB, CF::SH, kc, CF::RW * wave + 16 * mt);
        lds_ring_ld<2 * MT, kRing>([&](int i) { return frag_st(stA, CF::SH, i / MT, 16 * (i % MT)); },
                                   [&](int i, const bf16x8 b) {
                                     const int kc = i / MT, c = i % MT;
#pragma unroll
                                     for (int mt = 0; mt < MTW; ++mt) dw1[mt][c] = mfma(a[mt][kc], b, dw1[mt][c]);
                                   });
      }
      // dW0[p][f] += dZ0 X^T, X^T by transposing reads of the X rows (K
      // position p of the round at pass texel (p / 16) 16 NT + 16 rnd + p % 16
      // for W = 4, 64 rnd + p for NT = 1)
      {
        constexpr int NC = FP / 16;
        auto xrow = [&](int kc, int h) {
          const int p = 32 * kc + 8 * g + 4 * h + (n16 >> 2);
          const int t = kNT == 1 ? 64 * rnd + p : (p >> 4) * 16 * kNT + 16 * rnd + (p & 15);
          return t * CF::XS;
        };
        bf16x8 a[MTW][2];
#pragma unroll
        for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
          for (int kc = 0; kc < 2; ++kc) a[mt][kc] = frag_st(stC, CF::SH, kc, CF::RW * wave + 16 * mt);
        lds_ring_ld<2 * NC, kRing>(
            [&](int i) {
              const int c = 16 * (i % NC) + tq, kc = i / NC;
              return frag_tr(&xnf[xrow(kc, 0) + c], &xnf[xrow(kc, 1) + c]);
            },
            [&](int i, const bf16x8 b) {
              const int kc = i / NC, c = i % NC;
#pragma unroll
              for (int mt = 0; mt < MTW; ++mt) dw0[mt][c] = mfma(a[mt][kc], b, dw0[mt][c]);
            });
      }
      // dWout[o][p] += dY A1^T, column tiles of this wave
#pragma unroll
      for (int kc = 0; kc < 2; ++kc)
      {
        const bf16x8 a = frag_st(stY, 16, kc, 0);
#pragma unroll
        for (int c = 0; c < NTO; ++c) dwo[c] = mfma(a, frag_st(stD, CF::SH, kc, (wave * NTO + c) * 16), dwo[c]);
      }
      __syncthreads();
    }
    // the rounds ended with a barrier: X rows and target words are free
    x_commit();
    targets_commit(tg, tgt_s, wave, n16, g);
  }

  // partial decoder gradients of this workgroup: W0[H][F] | W1[H][H] | Wout[16][H]
  float* part = P.dw_part + (int64_t)vblk * P.dec_elems;
  const int F = P.F;
  float* pw0 = part;
  float* pw1 = part + H * F;
  float* pwo = pw1 + H * H;
#pragma unroll
  for (int mt = 0; mt < MTW; ++mt)
  {
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      const int row = kperm_pos(CF::RW * wave + 16 * mt + 4 * g + i);
#pragma unroll
      for (int c = 0; c < MT; ++c) pw1[row * H + kperm_pos(16 * c + n16)] = dw1[mt][c][i];
#pragma unroll
      for (int c = 0; c < FP / 16; ++c)
      {
        const int col = 16 * c + n16;
        if (col < F) pw0[row * F + col] = dw0[mt][c][i];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NTO; ++c)
#pragma unroll
    for (int i = 0; i < 4; ++i) pwo[(4 * g + i) * H + kperm_pos((wave * NTO + c) * 16 + n16)] = dwo[c][i];

  __syncthreads();
  double* red = reinterpret_cast<double*>(stA);
  red[tid] = sse;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1)
  {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) P.sse_part[vblk] = red[0];
}
} // namespace
