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
template <int H, int FP, int G, int ACT, int W, int FEAT>
struct MlpCfg
{
  static constexpr int NW = W;
  static constexpr int THREADS = 64 * W;
  static constexpr int NT = kSplitPass / (16 * W);
  static constexpr int PASS = kSplitPass;
  static constexpr int ROUNDS = PASS / 64;
  static constexpr int MT = H / 16, KF = FP / 32, KH = H / 32, GT = G / 16;
  static constexpr int RW = H / W, MTW = RW / 16, NTO = H / (16 * W);
  // dW1 (MT x MT tiles): each wave accumulates a BM1 x BN1 block of tiles
  // (fewer operand reads per round than a row of tiles); dW0: MTW rows of
  // tiles of its NCW column tiles that hold features (the padding columns
  // FEAT .. FP - 1 of the last tiles are never written out)
  static constexpr int T1 = MT * MT / W;
  static constexpr int BN1 = T1 >= 16 ? 4 : (T1 >= 4 ? 2 : 1);
  static constexpr int BM1 = T1 / BN1;
  static constexpr int NCW = (FEAT + 15) / 16;
  static_assert(NT >= 1 && MTW >= 1 && NTO >= 1 && (W == 4 || NT == 1), "wave shapes");
  static_assert(MT % BM1 == 0 && MT % BN1 == 0 && (MT / BM1) * (MT / BN1) == W, "dW1 blocks");
  static_assert(FEAT <= FP && NCW <= FP / 16, "feature columns");
  static constexpr int WT = H * FP + H * H + 16 * H;
  // X rows of 64 slots and staging rows of 128 units and more are unpadded,
  // their 16-byte chunks XOR-permuted by row (x_col, st_col); other widths
  // keep a padded stride
  static constexpr bool XSWZ = FP == 64;
  static constexpr bool SSWZ = H >= 128 && (H & (H - 1)) == 0;
  static constexpr int XS = XSWZ ? FP : FP + 8;  // LDS X row stride (elements)
  static constexpr int XR = PASS * XS;
  static constexpr int SH = SSWZ ? H : H + 16;   // staging row stride
  static constexpr int STAGE = 4 * 64 * SH + 64 * 16;
  static constexpr int LDS = WT + XR + STAGE;  // BF16 elements
  static_assert(2 * LDS + 16 * PASS <= 160 * 1024, "LDS budget");
  static_assert(STAGE * 2 >= 8 * 256, "final reduction reuses the staging region");
};

// Position in its row of column col of X row t (rows of 128 bytes): the
// chunk XOR key has bit 1 = t bit 1, bit 2 = t bit 2 ^ t bit 3, as for the
// 128-byte weight rows. The row reads of the layer-0 B operand (rows
// 16 w + i, chunk 4 j + g) put each 16-lane cycle of a 16-byte read on 16
// different chunks, and the transposing X^T reads of the decoder-gradient
// rounds (rows 8 g' + 4 h + k, both halves of chunk pairs) put each 32-lane
// cycle of an 8-byte read on 64 different banks.
template <bool kSwz>
__host__ __device__ __forceinline__ int x_col(int t, int col)
{
  if (!kSwz) return col;
  const int key = (t & 2) | ((((t >> 2) ^ (t >> 3)) & 1) << 2);
  return col ^ (key << 3);
}

// Position in its row of column col of staging row r (rows of 256 bytes or
// more): chunk XOR key 2 (r mod 4) + (r & 8). The wave writes of a round
// (rows 16 w + i, chunk 4 j + g; 16-byte writes go in eight cycles of lane
// quads {0, 2}, {1, 3}, {4, 6}, {5, 7} of each half wave) and the
// transposing reads of the operands (rows 8 g + k and 8 g + 4 + k, both
// halves of a chunk pair) are free of bank conflicts. Rows r and r + 4 share
// their key.
template <bool kSwz>
__host__ __device__ __forceinline__ int st_col(int r, int col)
{
  if (!kSwz) return col;
  return col ^ ((((r & 3) << 1) | (r & 8)) << 3);
}

// dY staging rows (16 BF16, 32 bytes): row r sits in slot r ^ 4 (r bit 3),
// so that the eight rows of a 32-lane transposing read (8 g + k) fill 256
// different bytes.
__host__ __device__ __forceinline__ int y_slot(int r) { return r ^ ((r >> 1) & 4); }

// Feature row stride in HBM: F rounded up to 8 elements (16-byte chunks).
__host__ __device__ constexpr int split_fxs(int F) { return (F + 7) & ~7; }

// A pass's feature rows (n 16-byte chunks) held in registers between their
// global loads and their LDS stores: up to three chunks per thread, as named
// members (an array here was kept in scratch memory by the compiler, which
// made every pass wait for its loads where they were issued).
struct XRows
{
  uint4 r0, r1, r2;
};

template <int kXL, int kThreads>
__device__ __forceinline__ XRows x_issue(const uint16_t* xg, int64_t pass, int pass_texels, int fxs, int n, int tid)
{
  static_assert(kXL >= 1 && kXL <= 3, "prefetch chunks per thread");
  const uint4* src = reinterpret_cast<const uint4*>(xg + pass * pass_texels * fxs);
  XRows x;
  const int c0 = tid, c1 = tid + kThreads, c2 = tid + 2 * kThreads;
  // unconditional loads from valid addresses
  x.r0 = src[c0 < n ? c0 : 0];
  x.r1 = kXL > 1 ? src[c1 < n ? c1 : 0] : make_uint4(0, 0, 0, 0);
  x.r2 = kXL > 2 ? src[c2 < n ? c2 : 0] : make_uint4(0, 0, 0, 0);
  return x;
}

template <int kXL, int kThreads, bool kSwz = false>
__device__ __forceinline__ void x_commit(const XRows& x, uint16_t* xnf, int xs, int xchunks, int n, int tid)
{
  auto put = [&](int c, const uint4& v) {
    if (c < n)
    {
      const int t = c / xchunks, q = c - t * xchunks;
      *reinterpret_cast<uint4*>(&xnf[t * xs + x_col<kSwz>(t, 8 * q)]) = v;
    }
  };
  put(tid, x.r0);
  if (kXL > 1) put(tid + kThreads, x.r1);
  if (kXL > 2) put(tid + 2 * kThreads, x.r2);
}

template <int H, int FP, int G, int ACT, int W, int RING, int FEAT>
__device__ __forceinline__ void mlp_body(const SplitParams& P, const int vblk, const int vgrid)
{
  using CF = MlpCfg<H, FP, G, ACT, W, FEAT>;
  constexpr int BM1 = CF::BM1, BN1 = CF::BN1, NCW = CF::NCW;
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
    {
      const int t = i / CF::XS;
      if (x_col<CF::XSWZ>(t, i - t * CF::XS) >= FXS) xnf[i] = 0;
    }
  }

  f32x4 dw1[BM1][BN1], dw0[MTW][NCW], dwo[NTO];
  const int mb1 = wave / (MT / BN1), nb1 = wave % (MT / BN1);  // this wave's dW1 block
#pragma unroll
  for (int a = 0; a < BM1; ++a)
#pragma unroll
    for (int c = 0; c < BN1; ++c) dw1[a][c] = f32x4{0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < MTW; ++a)
#pragma unroll
    for (int c = 0; c < NCW; ++c) dw0[a][c] = f32x4{0, 0, 0, 0};
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
  const int xn = kPass * xchunks;
  {
    uint32_t tg[kNT][4];
    const PassMap pm = pass_map((int64_t)vblk, T, kPass);
    const XRows xr = x_issue<kXL, kThreads>(P.xg, vblk, kPass, FXS, xn, tid);
    targets_issue<CF, kNT>(tg, pm, (int64_t)vblk, P, orig_s, wave, n16, g);
    x_commit<kXL, kThreads, CF::XSWZ>(xr, xnf, CF::XS, xchunks, xn, tid);
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
        for (int nt = 0; nt < kNT; ++nt)
          xb[j][nt] = frag_lds(&xnf[tloc[nt] * CF::XS + x_col<CF::XSWZ>(tloc[nt], 32 * j + 8 * g)]);
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
      for (int nt = 0; nt < kNT; ++nt) a0f[j][nt] = act_frag<ACT>(z0[2 * j][nt], z0[2 * j + 1][nt]);
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
      for (int nt = 0; nt < kNT; ++nt) a1f[j][nt] = act_frag<ACT>(z1[2 * j][nt], z1[2 * j + 1][nt]);
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
            const float yt = byte_over_255((tgt_s[tloc[nt] * 4 + g] >> (8 * i)) & 0xFFu);
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
            for (int nt = 0; nt < kNT; ++nt) dz1[r][nt] = mfma(a, dyf[nt], f32x4{0, 0, 0, 0});
          });
#pragma unroll
      for (int j = 0; j < KH; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
          dz1f[j][nt] = act_bwd_frag<ACT>(dz1[2 * j][nt], dz1[2 * j + 1][nt], a1f[j][nt], z1[2 * j][nt],
                                          z1[2 * j + 1][nt]);
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
          });
#pragma unroll
      for (int j = 0; j < KH; ++j)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
          dz0f[j][nt] = act_bwd_frag<ACT>(dz0[2 * j][nt], dz0[2 * j + 1][nt], a0f[j][nt], z0[2 * j][nt],
                                          z0[2 * j + 1][nt]);
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
    const XRows xr = x_issue<kXL, kThreads>(P.xg, pn, kPass, FXS, xn, tid);
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
          const int at = row * CF::SH + st_col<CF::SSWZ>(row, 32 * j + 8 * g);
          *reinterpret_cast<bf16x8*>(&stA[at]) = a0f[j][st];
          *reinterpret_cast<bf16x8*>(&stB[at]) = dz1f[j][st];
          *reinterpret_cast<bf16x8*>(&stC[at]) = dz0f[j][st];
          *reinterpret_cast<bf16x8*>(&stD[at]) = a1f[j][st];
        }
        *reinterpret_cast<uint2*>(&stY[y_slot(row) * 16 + 4 * g]) =
            make_uint2(pk_bf16(dy[st][0], dy[st][1]), pk_bf16(dy[st][2], dy[st][3]));
      }
      __syncthreads();
      const int trr0 = 8 * g + (n16 >> 2), tq = 4 * (n16 & 3);
      auto frag_st = [&](const uint16_t* st, int kc, int c0) {
        const int r = 32 * kc + trr0;
        const uint16_t* p = st + r * CF::SH + st_col<CF::SSWZ>(r, c0 + tq);
        return frag_tr(p, p + 4 * CF::SH);
      };
      auto frag_y = [&](int kc) {
        const int r = 32 * kc + trr0;
        return frag_tr(stY + y_slot(r) * 16 + tq, stY + y_slot(r + 4) * 16 + tq);
      };
      // dW1[p_out][p_in] += dZ1 A0^T over this wave's block of tiles; each
      // A0^T fragment read once and used for the block's row tiles
#pragma unroll
      for (int kc = 0; kc < 2; ++kc)
      {
        bf16x8 a[BM1];
#pragma unroll
        for (int mt = 0; mt < BM1; ++mt) a[mt] = frag_st(stB, kc, 16 * (BM1 * mb1 + mt));
        lds_ring_ld<BN1, kRing>([&](int c) { return frag_st(stA, kc, 16 * (BN1 * nb1 + c)); },
                                [&](int c, const bf16x8 b) {
#pragma unroll
                                  for (int mt = 0; mt < BM1; ++mt) dw1[mt][c] = mfma(a[mt], b, dw1[mt][c]);
                                });
      }
      // dW0[p][f] += dZ0 X^T, X^T by transposing reads of the X rows (K
      // position p of the round at pass texel (p / 16) 16 NT + 16 rnd + p % 16
      // for W = 4, 64 rnd + p for NT = 1)
      {
        constexpr int NC = NCW;
        auto xt = [&](int kc, int h, int c) {
          const int p = 32 * kc + 8 * g + 4 * h + (n16 >> 2);
          const int t = kNT == 1 ? 64 * rnd + p : (p >> 4) * 16 * kNT + 16 * rnd + (p & 15);
          return &xnf[t * CF::XS + x_col<CF::XSWZ>(t, c)];
        };
        bf16x8 a[MTW][2];
#pragma unroll
        for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
          for (int kc = 0; kc < 2; ++kc) a[mt][kc] = frag_st(stC, kc, CF::RW * wave + 16 * mt);
        lds_ring_ld<2 * NC, kRing>(
            [&](int i) {
              const int c = 16 * (i % NC) + tq, kc = i / NC;
              return frag_tr(xt(kc, 0, c), xt(kc, 1, c));
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
        const bf16x8 a = frag_y(kc);
#pragma unroll
        for (int c = 0; c < NTO; ++c) dwo[c] = mfma(a, frag_st(stD, kc, (wave * NTO + c) * 16), dwo[c]);
      }
      __syncthreads();
    }
    // the rounds ended with a barrier: X rows and target words are free
    x_commit<kXL, kThreads, CF::XSWZ>(xr, xnf, CF::XS, xchunks, xn, tid);
    targets_commit(tg, tgt_s, wave, n16, g);
  }

  // partial decoder gradients of this workgroup: W0[H][F] | W1[H][H] | Wout[16][H]
  float* part = P.dw_part + (int64_t)vblk * P.dec_elems;
  const int F = P.F;
  float* pw0 = part;
  float* pw1 = part + H * F;
  float* pwo = pw1 + H * H;
#pragma unroll
  for (int mt = 0; mt < BM1; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      const int row = kperm_pos(16 * (BM1 * mb1 + mt) + 4 * g + i);
#pragma unroll
      for (int c = 0; c < BN1; ++c) pw1[row * H + kperm_pos(16 * (BN1 * nb1 + c) + n16)] = dw1[mt][c][i];
    }
#pragma unroll
  for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      const int row = kperm_pos(CF::RW * wave + 16 * mt + 4 * g + i);
#pragma unroll
      for (int c = 0; c < NCW; ++c)
      {
        const int col = 16 * c + n16;
        if (col < F) pw0[row * F + col] = dw0[mt][c][i];
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
