/* Type-generic part of the oracle; included twice by ntc_oracle.c with
 * REAL = float / double. Test infrastructure only (see ntc_oracle.h).
 *
 * Requires macros: REAL, SFX(name), FMA, FLOOR, FABS. */

/* Gather: ntc_kernel.cu:75-177 (draw_features).
 * Per texel of tile b: position (x, y) = ((xy % T + ox_b) % M, (xy / T + oy_b) % M),
 * u = (x + 0.5) / M, grid coordinate g = u*D - 0.5, corners at floor(g)+{0,1}
 * clamped to [0, D-1]; corner c = j + 2k. g1 (and g0 in bilinear mode) slots
 * hold sum_c w_c * (G + N) accumulated with FMA from 0 in corner order
 * (w = {(1-fx)(1-fy), fx(1-fy), (1-fx)fy, fx fy}); g0 in corner mode holds the
 * 4 corners unweighted in slot c*c0 + ch (ntc_kernel.cu:119-143). Then the
 * triangle-wave encoding (ntc_kernel.cu:145-165): for each period P (first,
 * first/2, ...), axis, phase in {0, P/4}: 1 - 4|frac((coord+0.5+phase)/P) - 0.5|.
 * Then the LOD slot (ntc_kernel.cu:166-169) and zeros up to F (170-176). */
void SFX(o_draw_features)(int32_t g0_dim, int32_t g1_dim, int32_t c0,
                          int32_t c1, int32_t g0_mode, int32_t pe_first,
                          int32_t pe_count, int32_t lod_slot, int32_t F,
                          const int32_t* origins, int32_t n_tiles, int32_t T,
                          int32_t M, float norm_lod, uint32_t seed,
                          uint32_t step, int32_t level, int32_t noise_on,
                          float d0, float d1, const REAL* g0, const REAL* g1,
                          REAL* x)
{
  const int64_t TT = (int64_t)T * T;
  const int64_t B = (int64_t)n_tiles * TT;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < B; ++t)
  {
    const int32_t b = (int32_t)(t / TT);
    const int32_t xy = (int32_t)(t % TT);
    const int32_t px = (xy % T + origins[2 * b]) % M;
    const int32_t py = (xy / T + origins[2 * b + 1]) % M;
    const REAL u = ((REAL)px + (REAL)0.5) / (REAL)M;
    const REAL v = ((REAL)py + (REAL)0.5) / (REAL)M;
    const REAL g0x = u * (REAL)g0_dim - (REAL)0.5;
    const REAL g0y = v * (REAL)g0_dim - (REAL)0.5;
    const REAL g1x = u * (REAL)g1_dim - (REAL)0.5;
    const REAL g1y = v * (REAL)g1_dim - (REAL)0.5;
    const int32_t g0x0 = (int32_t)FLOOR(g0x), g0y0 = (int32_t)FLOOR(g0y);
    const int32_t g1x0 = (int32_t)FLOOR(g1x), g1y0 = (int32_t)FLOOR(g1y);
    const REAL f0x = g0x - FLOOR(g0x), f0y = g0y - FLOOR(g0y);
    const REAL f1x = g1x - FLOOR(g1x), f1y = g1y - FLOOR(g1y);
    const REAL one = (REAL)1;
    const REAL w0[4] = {(one - f0x) * (one - f0y), f0x * (one - f0y),
                        (one - f0x) * f0y, f0x * f0y};
    const REAL w1[4] = {(one - f1x) * (one - f1y), f1x * (one - f1y),
                        (one - f1x) * f1y, f1x * f1y};
    REAL acc0[64];
    REAL acc1[64];
    for (int ch = 0; ch < c0; ++ch) acc0[ch] = 0;
    for (int ch = 0; ch < c1; ++ch) acc1[ch] = 0;
    for (int c = 0; c < 4; ++c)
    {
      const int j = c & 1, k = c >> 1;
      const int32_t ax = o_clampi(g0x0 + j, 0, g0_dim - 1);
      const int32_t ay = o_clampi(g0y0 + k, 0, g0_dim - 1);
      const uint32_t cell0 = (uint32_t)(ay * g0_dim + ax);
      for (int ch = 0; ch < c0; ++ch)
      {
        REAL val = g0[(int64_t)cell0 * c0 + ch];
        if (noise_on) val = val + (REAL)o_noise(seed, step, 0, level, cell0, ch, d0);
        if (g0_mode == 0)
          x[(int64_t)(c * c0 + ch) * B + t] = val;
        else
          acc0[ch] = FMA(w0[c], val, acc0[ch]);
      }
      const int32_t bx = o_clampi(g1x0 + j, 0, g1_dim - 1);
      const int32_t by = o_clampi(g1y0 + k, 0, g1_dim - 1);
      const uint32_t cell1 = (uint32_t)(by * g1_dim + bx);
      for (int ch = 0; ch < c1; ++ch)
      {
        REAL val = g1[(int64_t)cell1 * c1 + ch];
        if (noise_on) val = val + (REAL)o_noise(seed, step, 1, level, cell1, ch, d1);
        acc1[ch] = FMA(w1[c], val, acc1[ch]);
      }
    }
    int slot = 0;
    if (g0_mode == 0)
      slot = 4 * c0;
    else
    {
      for (int ch = 0; ch < c0; ++ch) x[(int64_t)ch * B + t] = acc0[ch];
      slot = c0;
    }
    for (int ch = 0; ch < c1; ++ch) x[(int64_t)(slot + ch) * B + t] = acc1[ch];
    slot += c1;
    for (int p = 0; p < pe_count; ++p)
    {
      const int32_t P = pe_first >> p;
      for (int axis = 0; axis < 2; ++axis)
      {
        for (int h = 0; h < 2; ++h)
        {
          const REAL coord = (REAL)(axis == 0 ? px : py);
          const REAL phase = h == 0 ? (REAL)0 : (REAL)0.25 * (REAL)P;
          const REAL tt = (coord + (REAL)0.5 + phase) / (REAL)P;
          const REAL s = tt - FLOOR(tt);
          x[(int64_t)slot * B + t] = one - (REAL)4 * FABS(s - (REAL)0.5);
          ++slot;
        }
      }
    }
    if (lod_slot)
    {
      x[(int64_t)slot * B + t] = (REAL)norm_lod;
      ++slot;
    }
    for (; slot < F; ++slot) x[(int64_t)slot * B + t] = 0;
  }
}

/* Targets: ntc_kernel.cu:216-238 (draw_targets). */
void SFX(o_draw_targets)(const int32_t* origins, int32_t n_tiles, int32_t T,
                         int32_t M, int32_t O, const uint8_t* pkg, REAL* y)
{
  const int64_t TT = (int64_t)T * T;
  const int64_t B = (int64_t)n_tiles * TT;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < B; ++t)
  {
    const int32_t b = (int32_t)(t / TT);
    const int32_t xy = (int32_t)(t % TT);
    const int32_t px = (xy % T + origins[2 * b]) % M;
    const int32_t py = (xy / T + origins[2 * b + 1]) % M;
    for (int o = 0; o < O; ++o)
      y[(int64_t)o * B + t] = (REAL)pkg[((int64_t)py * M + px) * O + o] / (REAL)255;
  }
}

/* GEMMs: libopenntc.cpp:395-417 (matmulAB / matmulATB / matmulABT), row-major.
 * Each output element is one FMA chain in k order starting from 0; for ABT
 * (k = texels) the chain restarts every O_KCHUNK texels and the chunk sums are
 * added in order. */
void SFX(o_gemm)(int32_t op, int32_t n, int32_t m, int32_t k, const REAL* a,
                 const REAL* b, REAL* c)
{
  if (op == 0 || op == 1)
  {
    const int32_t JB = 512;
    const int64_t nblk = (int64_t)n * ((m + JB - 1) / JB);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nblk; ++blk)
    {
      const int32_t i = (int32_t)(blk % n);
      const int32_t j0 = (int32_t)(blk / n) * JB;
      const int32_t j1 = j0 + JB < m ? j0 + JB : m;
      REAL acc[512];
      for (int32_t j = j0; j < j1; ++j) acc[j - j0] = 0;
      for (int32_t kk = 0; kk < k; ++kk)
      {
        const REAL av = op == 0 ? a[(int64_t)i * k + kk] : a[(int64_t)kk * n + i];
        const REAL* brow = b + (int64_t)kk * m;
        for (int32_t j = j0; j < j1; ++j) acc[j - j0] = FMA(av, brow[j], acc[j - j0]);
      }
      for (int32_t j = j0; j < j1; ++j) c[(int64_t)i * m + j] = acc[j - j0];
    }
  }
  else
  {
    const int64_t pairs = (int64_t)n * m;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t pj = 0; pj < pairs; pj += 8)
    {
      const int32_t cnt = pairs - pj < 8 ? (int32_t)(pairs - pj) : 8;
      REAL total[8];
      const REAL* arow[8];
      const REAL* brow[8];
      for (int q = 0; q < cnt; ++q)
      {
        const int64_t e = pj + q;
        arow[q] = a + (e / m) * (int64_t)k;
        brow[q] = b + (e % m) * (int64_t)k;
        total[q] = 0;
      }
      for (int32_t k0 = 0; k0 < k; k0 += O_KCHUNK)
      {
        const int32_t k1 = k0 + O_KCHUNK < k ? k0 + O_KCHUNK : k;
        REAL acc[8];
        for (int q = 0; q < cnt; ++q) acc[q] = 0;
        for (int32_t kk = k0; kk < k1; ++kk)
          for (int q = 0; q < cnt; ++q) acc[q] = FMA(arow[q][kk], brow[q][kk], acc[q]);
        for (int q = 0; q < cnt; ++q) total[q] = total[q] + acc[q];
      }
      for (int q = 0; q < cnt; ++q) c[pj + q] = total[q];
    }
  }
}

/* HardGELU: ntc_kernel.cu:13-29; ReLU for the configuration layouts. */
void SFX(o_act_forward)(int32_t act, int64_t n, const REAL* z, REAL* a)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
  {
    const REAL v = z[i];
    if (act == 0)
      a[i] = v < (REAL)-1.5 ? (REAL)0 : (v < (REAL)1.5 ? (v / (REAL)3) * (v + (REAL)1.5) : v);
    else
      a[i] = v > (REAL)0 ? v : (REAL)0;
  }
}

void SFX(o_act_backward)(int32_t act, int64_t n, const REAL* z, const REAL* g,
                         REAL* out)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
  {
    const REAL v = z[i];
    if (act == 0)
      out[i] = v < (REAL)-1.5 ? (REAL)0
             : (v < (REAL)1.5 ? (((REAL)2 * v + (REAL)1.5) * g[i]) / (REAL)3 : g[i]);
    else
      out[i] = v > (REAL)0 ? g[i] : (REAL)0;
  }
}

/* Loss: libopenntc.cpp:659-668. r = y* - y^, dy = scale * r, SSE in double. */
double SFX(o_loss)(int64_t n, REAL scale, const REAL* yt, const REAL* yp,
                   REAL* dy)
{
  double sse = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sse)
  for (int64_t i = 0; i < n; ++i)
  {
    const REAL r = yt[i] - yp[i];
    if (dy) dy[i] = scale * r;
    sse += (double)r * (double)r;
  }
  return sse;
}

/* Scatter: ntc_kernel.cu:280-341 (accumulate_grid_gradients), the adjoint of
 * the grid part of the gather. Sequential in texel, corner, channel order. */
void SFX(o_scatter)(int32_t g0_dim, int32_t g1_dim, int32_t c0, int32_t c1,
                    int32_t g0_mode, const int32_t* origins, int32_t n_tiles,
                    int32_t T, int32_t M, const REAL* dx, REAL* dg0,
                    REAL* dg1)
{
  const int64_t TT = (int64_t)T * T;
  const int64_t B = (int64_t)n_tiles * TT;
  const int32_t g1_base = g0_mode == 0 ? 4 * c0 : c0;
  for (int64_t t = 0; t < B; ++t)
  {
    const int32_t b = (int32_t)(t / TT);
    const int32_t xy = (int32_t)(t % TT);
    const int32_t px = (xy % T + origins[2 * b]) % M;
    const int32_t py = (xy / T + origins[2 * b + 1]) % M;
    const REAL u = ((REAL)px + (REAL)0.5) / (REAL)M;
    const REAL v = ((REAL)py + (REAL)0.5) / (REAL)M;
    const REAL g0x = u * (REAL)g0_dim - (REAL)0.5;
    const REAL g0y = v * (REAL)g0_dim - (REAL)0.5;
    const REAL g1x = u * (REAL)g1_dim - (REAL)0.5;
    const REAL g1y = v * (REAL)g1_dim - (REAL)0.5;
    const int32_t g0x0 = (int32_t)FLOOR(g0x), g0y0 = (int32_t)FLOOR(g0y);
    const int32_t g1x0 = (int32_t)FLOOR(g1x), g1y0 = (int32_t)FLOOR(g1y);
    const REAL f0x = g0x - FLOOR(g0x), f0y = g0y - FLOOR(g0y);
    const REAL f1x = g1x - FLOOR(g1x), f1y = g1y - FLOOR(g1y);
    const REAL one = (REAL)1;
    const REAL w0[4] = {(one - f0x) * (one - f0y), f0x * (one - f0y),
                        (one - f0x) * f0y, f0x * f0y};
    const REAL w1[4] = {(one - f1x) * (one - f1y), f1x * (one - f1y),
                        (one - f1x) * f1y, f1x * f1y};
    for (int c = 0; c < 4; ++c)
    {
      const int j = c & 1, k = c >> 1;
      const int32_t ax = o_clampi(g0x0 + j, 0, g0_dim - 1);
      const int32_t ay = o_clampi(g0y0 + k, 0, g0_dim - 1);
      const int64_t cell0 = (int64_t)ay * g0_dim + ax;
      for (int ch = 0; ch < c0; ++ch)
      {
        if (g0_mode == 0)
          dg0[cell0 * c0 + ch] += dx[(int64_t)(c * c0 + ch) * B + t];
        else
          dg0[cell0 * c0 + ch] += w0[c] * dx[(int64_t)ch * B + t];
      }
      const int32_t bx = o_clampi(g1x0 + j, 0, g1_dim - 1);
      const int32_t by = o_clampi(g1y0 + k, 0, g1_dim - 1);
      const int64_t cell1 = (int64_t)by * g1_dim + bx;
      for (int ch = 0; ch < c1; ++ch)
        dg1[cell1 * c1 + ch] += w1[c] * dx[(int64_t)(g1_base + ch) * B + t];
    }
  }
}

/* Decoder forward, loss and backward: libopenntc.cpp:651-693.
 * ws: Z0 A0 Z1 A1 (H*B each), Yp dY (O*B each), dA1 dZ1 dA0 dZ0 (H*B each). */
double SFX(o_mlp_fwd_bwd)(int32_t F, int32_t H, int32_t O, int32_t act,
                          int64_t B, const REAL* x, const REAL* yt,
                          const REAL* w0, const REAL* w1, const REAL* wout,
                          REAL* dw0, REAL* dw1, REAL* dwout, REAL* dx,
                          REAL* ws)
{
  const int32_t Bi = (int32_t)B;
  REAL* z0 = ws;
  REAL* a0 = z0 + (int64_t)H * B;
  REAL* z1 = a0 + (int64_t)H * B;
  REAL* a1 = z1 + (int64_t)H * B;
  REAL* yp = a1 + (int64_t)H * B;
  REAL* dy = yp + (int64_t)O * B;
  REAL* da1 = dy + (int64_t)O * B;
  REAL* dz1 = da1 + (int64_t)H * B;
  REAL* da0 = dz1 + (int64_t)H * B;
  REAL* dz0 = da0 + (int64_t)H * B;
  SFX(o_gemm)(0, H, Bi, F, w0, x, z0);
  SFX(o_act_forward)(act, (int64_t)H * B, z0, a0);
  SFX(o_gemm)(0, H, Bi, H, w1, a0, z1);
  SFX(o_act_forward)(act, (int64_t)H * B, z1, a1);
  SFX(o_gemm)(0, O, Bi, H, wout, a1, yp);
  /* -2.0f / ((float)out_dim_ * batch_dim) in the working precision */
  const REAL scale = (REAL)-2 / ((REAL)O * (REAL)B);
  const double sse = SFX(o_loss)((int64_t)O * B, scale, yt, yp, dy);
  SFX(o_gemm)(2, O, H, Bi, dy, a1, dwout);
  SFX(o_gemm)(1, H, Bi, O, wout, dy, da1);
  SFX(o_act_backward)(act, (int64_t)H * B, z1, da1, dz1);
  SFX(o_gemm)(2, H, H, Bi, dz1, a0, dw1);
  SFX(o_gemm)(1, H, Bi, H, w1, dz1, da0);
  SFX(o_act_backward)(act, (int64_t)H * B, z0, da0, dz0);
  SFX(o_gemm)(2, H, F, Bi, dz0, x, dw0);
  SFX(o_gemm)(1, F, Bi, H, w0, dz0, dx);
  return sse;
}

double SFX(o_mlp_forward)(int32_t F, int32_t H, int32_t O, int32_t act,
                          int64_t B, const REAL* x, const REAL* yt,
                          const REAL* w0, const REAL* w1, const REAL* wout,
                          REAL* ws, REAL* max_a0, REAL* max_a1)
{
  const int32_t Bi = (int32_t)B;
  REAL* z0 = ws;
  REAL* a0 = z0 + (int64_t)H * B;
  REAL* z1 = a0 + (int64_t)H * B;
  REAL* a1 = z1 + (int64_t)H * B;
  REAL* yp = a1 + (int64_t)H * B;
  SFX(o_gemm)(0, H, Bi, F, w0, x, z0);
  SFX(o_act_forward)(act, (int64_t)H * B, z0, a0);
  SFX(o_gemm)(0, H, Bi, H, w1, a0, z1);
  SFX(o_act_forward)(act, (int64_t)H * B, z1, a1);
  if (max_a0)
  {
    REAL mx = *max_a0;
    for (int64_t i = 0; i < (int64_t)H * B; ++i) mx = FABS(a0[i]) > mx ? FABS(a0[i]) : mx;
    *max_a0 = mx;
  }
  if (max_a1)
  {
    REAL mx = *max_a1;
    for (int64_t i = 0; i < (int64_t)H * B; ++i) mx = FABS(a1[i]) > mx ? FABS(a1[i]) : mx;
    *max_a1 = mx;
  }
  if (!yt) return 0.0;
  SFX(o_gemm)(0, O, Bi, H, wout, a1, yp);
  return SFX(o_loss)((int64_t)O * B, (REAL)0, yt, yp, (REAL*)0);
}
