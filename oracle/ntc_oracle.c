/* ntc_oracle.c - CPU restatement of the openntc training hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels, loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. Never
 * linked into the product.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 *
 * Where the reference leaves a choice to a library we cannot run here, this
 * file defines it and the GPU path follows the same definition:
 *   - host RNG: std::mt19937 (the standard fixes it) with distributions
 *     defined below (the reference's MSVC <random> distributions are
 *     implementation-defined, libopenntc.cpp:564-583);
 *   - quantisation noise: Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
 *     "Parallel random numbers: as easy as 1, 2, 3", SC'11) instead of the
 *     cuRAND XORWOW stream (ntc_kernel.cu:412-446);
 *   - GEMMs: one FMA chain per output in k order (cuBLAS TF32 order is
 *     unknowable, libopenntc.cpp:374).
 *
 * Parity pins (tests/test_oracle_*.py): mt19937 10000th-output KAT, Philox
 * Random123 KATs, README size table, closed-form gather cases and the
 * gather/scatter adjoint identity, finite differences of the FP64 loss,
 * brute-force GEMMs, exhaustive pack/unpack, Lanczos invariants.
 */
#include "ntc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define O_KCHUNK 1024

static inline int32_t o_clampi(int32_t x, int32_t a, int32_t b)
{
  return x > b ? b : (x < a ? a : x);
}

/* ---- mt19937 (ISO C++ [rand.eng.mers], parameters of std::mt19937) ------- */
size_t o_mt_size(void) { return sizeof(o_mt); }

void o_mt_seed(o_mt* m, uint32_t seed)
{
  m->s[0] = seed;
  for (int i = 1; i < 624; ++i)
    m->s[i] = 1812433253u * (m->s[i - 1] ^ (m->s[i - 1] >> 30)) + (uint32_t)i;
  m->i = 624;
}

uint32_t o_mt_next(o_mt* m)
{
  if (m->i >= 624)
  {
    for (int k = 0; k < 624; ++k)
    {
      const uint32_t y = (m->s[k] & 0x80000000u) | (m->s[(k + 1) % 624] & 0x7fffffffu);
      m->s[k] = m->s[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    m->i = 0;
  }
  uint32_t y = m->s[m->i++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* Distributions used by the schedule and the parameter init. */
float o_uniform01(o_mt* m) { return (float)(o_mt_next(m) >> 8) * 0x1p-24f; }

float o_uniform01_open0(o_mt* m)
{
  return (float)((o_mt_next(m) >> 8) + 1u) * 0x1p-24f;
}

float o_uniform_range(o_mt* m, float lo, float hi)
{
  const float u = o_uniform01(m);
  return lo + (hi - lo) * u;
}

int32_t o_uniform_int(o_mt* m, int32_t lo, int32_t hi)
{
  const uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo) + 1u;
  const uint64_t limit = (((uint64_t)1 << 32) / range) * range;
  uint32_t u;
  do
  {
    u = o_mt_next(m);
  } while ((uint64_t)u >= limit);
  return lo + (int32_t)((uint64_t)u % range);
}

int32_t o_bernoulli(o_mt* m, double p)
{
  return ((double)o_mt_next(m) * 0x1p-32) < p ? 1 : 0;
}

/* ---- layout ----------------------------------------------------------------
 * Profiles: libopenntc.cpp:289-330. Quantisation step 2/2^b: :329-330.
 * F = roundup4(4*c0 + c1 + 12 + 1) (:332); blob pad roundup16(F) (:333).
 * Mips while > mip_min_dim (:1209-1216, reference stops at 4).
 * Levels: g0 = dim/s, g1 = g0/2, then /4 while g0/4 >= 4 (:1218-1227). */
static int o_layout_finish(o_layout* L, int32_t dim, int32_t out_dim,
                           int32_t mip_min_dim)
{
  if (dim < 2 || (dim & (dim - 1)) != 0 || out_dim < 1 || out_dim > 16) return -1;
  if (mip_min_dim <= 0) mip_min_dim = 4;
  L->dim = dim;
  L->out_dim = out_dim;
  L->out_dim_padded = 16;
  L->mip_dim[0] = dim;
  int n = 0;
  while (L->mip_dim[n] > mip_min_dim && n + 1 < 12)
  {
    L->mip_dim[n + 1] = L->mip_dim[n] / 2;
    ++n;
  }
  L->mip_count = n + 1;
  L->g0_dim[0] = dim / L->g0_scale;
  L->g1_dim[0] = L->g0_dim[0] / 2;
  if (L->g1_dim[0] < 1) return -1;
  int lv = 0;
  while (L->g0_dim[lv] / 4 >= 4 && lv + 1 < 6)
  {
    L->g0_dim[lv + 1] = L->g0_dim[lv] / 4;
    L->g1_dim[lv + 1] = L->g1_dim[lv] / 4;
    ++lv;
  }
  L->level_count = lv + 1;
  L->g0_delta = 2.0f / powf(2.0f, (float)L->g0_bits);
  L->g1_delta = 2.0f / powf(2.0f, (float)L->g1_bits);
  if (L->g0_mode == 0)
  {
    const int f = 4 * L->g0_channels + L->g1_channels + 4 * L->pe_count + L->lod_slot;
    L->feature_dim = ((f + 3) / 4) * 4;
  }
  else
    L->feature_dim = L->g0_channels + L->g1_channels + 4 * L->pe_count + L->lod_slot;
  L->feature_dim_padded = ((L->feature_dim + 15) / 16) * 16;
  return 0;
}

int o_layout_profile(int32_t profile, int32_t dim, int32_t out_dim,
                     int32_t mip_min_dim, o_layout* L)
{
  static const int32_t tab[4][5] = {
      /* b0 b1 c0 c1 s */
      {2, 4, 8, 12, 4},
      {4, 4, 12, 20, 4},
      {2, 4, 12, 10, 2},
      {4, 4, 16, 12, 2},
  };
  if (profile < 0 || profile > 3) return -1;
  memset(L, 0, sizeof(*L));
  L->g0_bits = tab[profile][0];
  L->g1_bits = tab[profile][1];
  L->g0_channels = tab[profile][2];
  L->g1_channels = tab[profile][3];
  L->g0_scale = tab[profile][4];
  L->hidden_dim = 64;
  L->g0_mode = 0;
  L->pe_first = 8;
  L->pe_count = 3;
  L->lod_slot = 1;
  L->activation = 0;
  return o_layout_finish(L, dim, out_dim, mip_min_dim);
}

int o_layout_custom(int32_t g0_bits, int32_t g1_bits, int32_t g0_ch,
                    int32_t g1_ch, int32_t g0_scale, int32_t hidden,
                    int32_t feature_layout, int32_t activation, int32_t dim,
                    int32_t out_dim, int32_t mip_min_dim, o_layout* L)
{
  memset(L, 0, sizeof(*L));
  L->g0_bits = g0_bits;
  L->g1_bits = g1_bits;
  L->g0_channels = g0_ch;
  L->g1_channels = g1_ch;
  L->g0_scale = g0_scale;
  L->hidden_dim = hidden;
  L->activation = activation;
  if (feature_layout == 0)
  {
    L->g0_mode = 0;
    L->pe_first = 8;
    L->pe_count = 3;
    L->lod_slot = 1;
  }
  else
  {
    L->g0_mode = 1;
    L->pe_first = 4;
    L->pe_count = 2;
    L->lod_slot = 0;
  }
  return o_layout_finish(L, dim, out_dim, mip_min_dim);
}

/* libopenntc.cpp:881-891 */
int32_t o_level_for_lod(const o_layout* L, int32_t lod)
{
  const int32_t first = (int32_t)log2f((float)L->g0_scale) + 2;
  if (lod < first) return 0;
  const int32_t level = 1 + (lod - first) / 2;
  return level < L->level_count - 1 ? level : L->level_count - 1;
}

/* libopenntc.cpp:381-387 (blob) and 1055-1074 (packed level sizes) */
uint64_t o_decoder_bytes(const o_layout* L)
{
  const uint64_t H = (uint64_t)L->hidden_dim;
  return H * L->feature_dim_padded + H * H + 16 * H + 4 * (H + H + 16);
}

uint64_t o_compressed_bytes(const o_layout* L)
{
  uint64_t total = 0;
  for (int l = 0; l < L->level_count; ++l)
  {
    total += (uint64_t)L->g0_dim[l] * L->g0_dim[l] * L->g0_channels * L->g0_bits / 8;
    total += (uint64_t)L->g1_dim[l] * L->g1_dim[l] * L->g1_channels * L->g1_bits / 8;
  }
  return total + o_decoder_bytes(L);
}

/* ---- init (libopenntc.cpp:60-78, 132-152, 483-501) ------------------------ */
void o_init_params(o_mt* m, const o_layout* L, float lo, float hi, float** g0,
                   float** g1, float* w0, float* w1, float* wout)
{
  for (int l = 0; l < L->level_count; ++l)
  {
    const int64_t n0 = (int64_t)L->g0_dim[l] * L->g0_dim[l] * L->g0_channels;
    for (int64_t i = 0; i < n0; ++i) g0[l][i] = o_uniform_range(m, lo, hi);
    const int64_t n1 = (int64_t)L->g1_dim[l] * L->g1_dim[l] * L->g1_channels;
    for (int64_t i = 0; i < n1; ++i) g1[l][i] = o_uniform_range(m, lo, hi);
  }
  const int32_t H = L->hidden_dim, F = L->feature_dim;
  float bound = 1.0f / sqrtf((float)F);
  for (int64_t i = 0; i < (int64_t)H * F; ++i) w0[i] = o_uniform_range(m, -bound, bound);
  bound = 1.0f / sqrtf((float)H);
  for (int64_t i = 0; i < (int64_t)H * H; ++i) w1[i] = o_uniform_range(m, -bound, bound);
  for (int64_t i = 0; i < (int64_t)16 * H; ++i) wout[i] = o_uniform_range(m, -bound, bound);
}

/* ---- schedule (libopenntc.cpp:419-423, 475, 564-585, 711-720) ------------ */
float o_cosine_annealing(float lr_min, float lr_max, int32_t t_max, int32_t t_cur)
{
  return lr_min + 0.5f * (lr_max - lr_min) *
                      (1.0f + cosf((float)t_cur * acosf(-1.0f) / (float)t_max));
}

void o_schedule(o_mt* m, const o_layout* L, int32_t n_steps,
                int32_t grids_per_batch, float lr_grid_max, float lr_dec_max,
                o_step* out)
{
  const int32_t lock = (95 * n_steps) / 100;
  int32_t counters[6] = {0, 0, 0, 0, 0, 0};
  for (int32_t t = 0; t < n_steps; ++t)
  {
    o_step* s = &out[t];
    memset(s, 0, sizeof(*s));
    const int32_t draw_uniform = o_bernoulli(m, 0.05);
    const float U = o_uniform01_open0(m);
    int32_t lod = (int32_t)floorf(0.5f * -log2f(U));
    if (lod > L->mip_count - 1) lod = L->mip_count - 1;
    const int32_t lod_uniform = o_uniform_int(m, 0, L->mip_count - 1);
    if (draw_uniform) lod = lod_uniform;
    s->lod = lod;
    s->level = o_level_for_lod(L, lod);
    s->n_tiles = grids_per_batch;
    const int32_t M = L->mip_dim[lod];
    s->tile_dim = M < 256 ? M : 256;
    for (int i = 0; i < grids_per_batch; ++i)
    {
      s->origins[2 * i] = o_uniform_int(m, 0, M - 1);
      s->origins[2 * i + 1] = o_uniform_int(m, 0, M - 1);
    }
    s->bias1 = 1.0f / (1.0f - powf(0.9f, (float)t + 1.0f));
    s->bias2 = 1.0f / (1.0f - powf(0.999f, (float)t + 1.0f));
    const int32_t gc = counters[s->level];
    s->grid_bias1 = 1.0f / (1.0f - powf(0.9f, (float)gc + 1.0f));
    s->grid_bias2 = 1.0f / (1.0f - powf(0.999f, (float)gc + 1.0f));
    counters[s->level] += 1;
    s->lr_grid = o_cosine_annealing(0.0f, lr_grid_max, n_steps, t);
    s->lr_decoder = o_cosine_annealing(0.0f, lr_dec_max, n_steps, t);
    s->grid_update = t < lock;
    s->lock = t == lock;
  }
}

/* ---- Philox4x32-10 ---------------------------------------------------------
 * Round: (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2;
 *        c = {hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0}; key += (W0, W1) between
 * rounds. Constants from the SC'11 paper. */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r)
  {
    if (r > 0)
    {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Noise of one grid element: U(0,1] * delta - delta/2 (ntc_kernel.cu:438),
 * U from Philox with counter {cell, ch/4, step, grid | level<<1} and key
 * {seed, 0x4E5443}; word ch%4. Exact in float (delta is a power of two). */
float o_noise(uint32_t seed, uint32_t step, int32_t grid, int32_t level,
              uint32_t cell, int32_t ch, float delta)
{
  const uint32_t ctr[4] = {cell, (uint32_t)ch >> 2, step,
                           (uint32_t)grid | ((uint32_t)level << 1)};
  const uint32_t key[2] = {seed, 0x4E5443u};
  uint32_t r[4];
  o_philox4x32_10(ctr, key, r);
  const float u = (float)((r[ch & 3] >> 8) + 1u) * 0x1p-24f;
  return u * delta - 0.5f * delta;
}

/* ---- generic part, float and double --------------------------------------- */
#define REAL float
#define SFX(name) name##_f32
#define FMA fmaf
#define FLOOR floorf
#define FABS fabsf
#include "ntc_oracle_impl.h"
#undef REAL
#undef SFX
#undef FMA
#undef FLOOR
#undef FABS

#define REAL double
#define SFX(name) name##_f64
#define FMA fma
#define FLOOR floor
#define FABS fabs
#include "ntc_oracle_impl.h"
#undef REAL
#undef SFX
#undef FMA
#undef FLOOR
#undef FABS

/* ---- optimiser and grid maintenance ---------------------------------------- */
/* ntc_kernel.cu:375-394: m = b1 m + (1-b1) g; v = b2 v + ((1-b2) g) g;
 * p -= (lr (m bias1)) / (sqrt(v bias2) + 1e-8). The two moment updates are
 * single FMAs (what nvcc emits for them). */
void o_adam_f32(int64_t n, float lr, float b1, float b2, float bias1,
                float bias2, const float* g, float* m, float* v, float* p)
{
  const float omb1 = 1.0f - b1, omb2 = 1.0f - b2;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
  {
    const float gi = g[i];
    const float mi = fmaf(b1, m[i], omb1 * gi);
    const float vi = fmaf(b2, v[i], (omb2 * gi) * gi);
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - (lr * (mi * bias1)) / (sqrtf(vi * bias2) + 1e-8f);
  }
}

/* ntc_kernel.cu:468-476 */
void o_clamp_grid_f32(int64_t n, float delta, float* g)
{
  const float hi = 1.0f - delta;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) g[i] = fmaxf(-1.0f, fminf(hi, g[i]));
}

/* ntc_kernel.cu:448-460 */
void o_quantize_grid_f32(int64_t n, int32_t bits, float delta, float* g)
{
  const float kmin = -(float)(1 << (bits - 1));
  const float kmax = (float)(1 << (bits - 1)) - 1.0f;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
  {
    const float k = rintf(g[i] / delta);
    g[i] = fminf(fmaxf(k, kmin), kmax) * delta;
  }
}

/* ntc_kernel.cu:486-499: word w holds elements w*(32/b)+s at bits s*b,
 * code = clamp(roundf(g * 2^(b-1))) + 2^(b-1). */
void o_quantize_pack_f32(int64_t n, int32_t bits, const float* in, uint32_t* out)
{
  const int32_t per = 32 / bits;
  const int64_t words = n * bits / 32;
  const int32_t half = 1 << (bits - 1);
  for (int64_t w = 0; w < words; ++w)
  {
    uint32_t packed = 0;
    for (int s = 0; s < per; ++s)
    {
      int32_t q = (int32_t)roundf(in[w * per + s] * (float)half);
      q = o_clampi(q, -half, half - 1);
      packed |= (uint32_t)(q + half) << (s * bits);
    }
    out[w] = packed;
  }
}

/* libopenntc.cpp:445-469 */
void o_quantize_weights(const float* w, int32_t rows, int32_t cols,
                        float prescale, uint32_t* wq, float* ws)
{
  for (int32_t i = 0; i < rows; ++i)
  {
    float max_abs = 0.0f;
    for (int32_t j = 0; j < cols; ++j)
    {
      const float a = fabsf(w[(int64_t)i * cols + j]);
      max_abs = a > max_abs ? a : max_abs;
    }
    const float scale = max_abs == 0.0f ? 1.0f : max_abs / 127.0f;
    ws[i] = scale * prescale;
    for (int32_t j = 0; j < cols; j += 4)
    {
      uint32_t pack = 0;
      for (int c = 0; c < 4; ++c)
      {
        const float val = w[(int64_t)i * cols + j + c] / scale;
        const int32_t q = o_clampi((int32_t)lroundf(val), -128, 127);
        pack |= ((uint32_t)q & 0xFFu) << (c * 8);
      }
      wq[(int64_t)i * (cols / 4) + j / 4] = pack;
    }
  }
}

/* ---- Lanczos-3 2:1 (ntc_kernel.cu:569-663) --------------------------------
 * Output texel t samples source 2t+1+i, i in [-6, 5], weight L((i+0.5)/2),
 * L(x) = sinc(pi x) sinc(pi x / 3); out-of-range taps skipped and the sum
 * renormalised. The 12 weights are evaluated in double and rounded to float
 * once. */
void o_lanczos_weights(float w[12])
{
  const double pi = 3.14159265358979323846;
  for (int i = -6; i < 6; ++i)
  {
    const double x = (0.5 + i) * 0.5;
    const double a = pi * x, b = pi * x / 3.0;
    const double sa = fabs(a) < 1e-8 ? 1.0 : sin(a) / a;
    const double sb = fabs(b) < 1e-8 ? 1.0 : sin(b) / b;
    w[i + 6] = (float)(sa * sb);
  }
}

void o_filter_lanczos(int32_t dim, int32_t c, const uint8_t* src, uint8_t* dst)
{
  float w[12];
  o_lanczos_weights(w);
  const int32_t half = dim / 2;
  float* tmp = (float*)malloc(sizeof(float) * (size_t)half * dim * c);
  /* pass 1: along x, uint8/255 -> float, [dim rows][half cols][c] */
#pragma omp parallel for schedule(static)
  for (int32_t y = 0; y < dim; ++y)
  {
    for (int32_t ox = 0; ox < half; ++ox)
    {
      float tot[4] = {0, 0, 0, 0};
      float tw = 0.0f;
      for (int i = -6; i < 6; ++i)
      {
        const int32_t s = 2 * ox + 1 + i;
        if (s < 0 || s >= dim) continue;
        const float wt = w[i + 6];
        for (int ch = 0; ch < c; ++ch)
          tot[ch] = fmaf((float)src[((int64_t)y * dim + s) * c + ch] / 255.0f, wt, tot[ch]);
        tw = tw + wt;
      }
      for (int ch = 0; ch < c; ++ch) tmp[((int64_t)y * half + ox) * c + ch] = tot[ch] / tw;
    }
  }
  /* pass 2: along y, float -> truncated uint8 */
#pragma omp parallel for schedule(static)
  for (int32_t oy = 0; oy < half; ++oy)
  {
    for (int32_t ox = 0; ox < half; ++ox)
    {
      float tot[4] = {0, 0, 0, 0};
      float tw = 0.0f;
      for (int i = -6; i < 6; ++i)
      {
        const int32_t s = 2 * oy + 1 + i;
        if (s < 0 || s >= dim) continue;
        const float wt = w[i + 6];
        for (int ch = 0; ch < c; ++ch)
          tot[ch] = fmaf(tmp[((int64_t)s * half + ox) * c + ch], wt, tot[ch]);
        tw = tw + wt;
      }
      for (int ch = 0; ch < c; ++ch)
      {
        const float val = tot[ch] / tw;
        const float q = fminf(255.0f, fmaxf(0.0f, val * 255.0f));
        dst[((int64_t)oy * half + ox) * c + ch] = (uint8_t)q;
      }
    }
  }
  free(tmp);
}

/* ---- full step ------------------------------------------------------------ */
static int64_t o_max_batch(const o_layout* L, int64_t max_batch)
{
  (void)L;
  return max_batch;
}

size_t o_step_workspace_elems(const o_layout* L, int64_t max_batch)
{
  const int64_t B = o_max_batch(L, max_batch);
  const int64_t F = L->feature_dim, H = L->hidden_dim, O = L->out_dim;
  const int64_t dec = H * F + H * H + 16 * H;
  const int64_t lvl = (int64_t)L->g0_dim[0] * L->g0_dim[0] * L->g0_channels +
                      (int64_t)L->g1_dim[0] * L->g1_dim[0] * L->g1_channels;
  /* X, dX, Yt, MLP workspace, decoder grads, level grads */
  return (size_t)(2 * F * B + O * B + 8 * H * B + 2 * O * B + dec + lvl);
}

void o_lock_grids_f32(const o_layout* L, float** g0, float** g1)
{
  for (int l = 0; l < L->level_count; ++l)
  {
    o_quantize_grid_f32((int64_t)L->g0_dim[l] * L->g0_dim[l] * L->g0_channels,
                        L->g0_bits, L->g0_delta, g0[l]);
    o_quantize_grid_f32((int64_t)L->g1_dim[l] * L->g1_dim[l] * L->g1_channels,
                        L->g1_bits, L->g1_delta, g1[l]);
  }
}

double o_train_step_f32(const o_layout* L, const o_step* st, uint32_t seed,
                        uint32_t step_index, float** g0, float** g1,
                        float** g0m, float** g0v, float** g1m, float** g1v,
                        float* dec, float* dec_m, float* dec_v,
                        const uint8_t** packages, float* ws)
{
  const int32_t F = L->feature_dim, H = L->hidden_dim, O = L->out_dim;
  const int32_t lod = st->lod, lvl = st->level, T = st->tile_dim;
  const int32_t M = L->mip_dim[lod];
  const int64_t B = (int64_t)st->n_tiles * T * T;
  const int32_t D0 = L->g0_dim[lvl], D1 = L->g1_dim[lvl];
  const int64_t n0 = (int64_t)D0 * D0 * L->g0_channels;
  const int64_t n1 = (int64_t)D1 * D1 * L->g1_channels;
  float* x = ws;
  float* dx = x + (int64_t)F * B;
  float* yt = dx + (int64_t)F * B;
  float* mlp = yt + (int64_t)O * B;
  float* dgrad = mlp + 8 * (int64_t)H * B + 2 * (int64_t)O * B;
  float* dg0 = dgrad + (int64_t)H * F + (int64_t)H * H + 16 * (int64_t)H;
  float* dg1 = dg0 + n0;

  if (st->lock) o_lock_grids_f32(L, g0, g1);
  const float norm_lod = (float)lod / (float)(L->mip_count - 1);
  o_draw_features_f32(D0, D1, L->g0_channels, L->g1_channels, L->g0_mode,
                      L->pe_first, L->pe_count, L->lod_slot, F, st->origins,
                      st->n_tiles, T, M, norm_lod, seed, step_index, lvl,
                      st->grid_update, L->g0_delta, L->g1_delta, g0[lvl],
                      g1[lvl], x);
  o_draw_targets_f32(st->origins, st->n_tiles, T, M, O, packages[lod], yt);
  float* w0 = dec;
  float* w1 = w0 + (int64_t)H * F;
  float* wout = w1 + (int64_t)H * H;
  float* dw0 = dgrad;
  float* dw1 = dw0 + (int64_t)H * F;
  float* dwout = dw1 + (int64_t)H * H;
  memset(dwout, 0, sizeof(float) * 16 * (size_t)H);
  const double sse = o_mlp_fwd_bwd_f32(F, H, O, L->activation, B, x, yt, w0,
                                       w1, wout, dw0, dw1, dwout, dx, mlp);
  if (st->grid_update)
  {
    memset(dg0, 0, sizeof(float) * (size_t)n0);
    memset(dg1, 0, sizeof(float) * (size_t)n1);
    o_scatter_f32(D0, D1, L->g0_channels, L->g1_channels, L->g0_mode,
                  st->origins, st->n_tiles, T, M, dx, dg0, dg1);
  }
  /* decoder Adam over W0 | W1 | Wout rows < O (libopenntc.cpp:722-754) */
  const int64_t ndec = (int64_t)H * F + (int64_t)H * H + (int64_t)O * H;
  o_adam_f32(ndec, st->lr_decoder, 0.9f, 0.999f, st->bias1, st->bias2, dgrad,
             dec_m, dec_v, dec);
  if (st->grid_update)
  {
    o_adam_f32(n0, st->lr_grid, 0.9f, 0.999f, st->grid_bias1, st->grid_bias2,
               dg0, g0m[lvl], g0v[lvl], g0[lvl]);
    o_adam_f32(n1, st->lr_grid, 0.9f, 0.999f, st->grid_bias1, st->grid_bias2,
               dg1, g1m[lvl], g1v[lvl], g1[lvl]);
    o_clamp_grid_f32(n0, L->g0_delta, g0[lvl]);
    o_clamp_grid_f32(n1, L->g1_delta, g1[lvl]);
  }
  return sse;
}

/* Tile walk shared by eval and calibration (libopenntc.cpp:912-932, 982-1005):
 * per mip, 256^2 tiles (or one mip-sized tile) in row-major order, <= 8 per
 * batch. */
typedef void (*o_batch_fn)(const o_layout* L, int32_t mip, const int32_t* origins,
                           int32_t n_tiles, int32_t T, void* user);

static void o_walk_tiles(const o_layout* L, o_batch_fn fn, void* user)
{
  for (int32_t mip = 0; mip < L->mip_count; ++mip)
  {
    const int32_t M = L->mip_dim[mip];
    const int32_t per_dim = (M + 255) / 256;
    const int32_t n_grids = per_dim * per_dim;
    const int32_t T = M < 256 ? M : 256;
    for (int32_t start = 0; start < n_grids; start += 8)
    {
      const int32_t nb = n_grids - start < 8 ? n_grids - start : 8;
      int32_t origins[16];
      for (int32_t g = 0; g < nb; ++g)
      {
        origins[2 * g] = ((start + g) % per_dim) * T;
        origins[2 * g + 1] = ((start + g) / per_dim) * T;
      }
      fn(L, mip, origins, nb, T, user);
    }
  }
}

typedef struct o_eval_ctx
{
  float** g0;
  float** g1;
  const float* dec;
  const uint8_t** packages;
  float* ws;
  double sse, count;
  double* channel_sse;
  float max_a0, max_a1;
  int calibrate;
} o_eval_ctx;

static void o_eval_batch(const o_layout* L, int32_t mip, const int32_t* origins,
                         int32_t n_tiles, int32_t T, void* user)
{
  o_eval_ctx* e = (o_eval_ctx*)user;
  const int32_t F = L->feature_dim, H = L->hidden_dim, O = L->out_dim;
  const int64_t B = (int64_t)n_tiles * T * T;
  const int32_t lvl = o_level_for_lod(L, mip);
  float* x = e->ws;
  float* yt = x + (int64_t)F * B;
  float* mlp = yt + (int64_t)O * B;
  const float norm_lod = (float)mip / (float)(L->mip_count - 1);
  o_draw_features_f32(L->g0_dim[lvl], L->g1_dim[lvl], L->g0_channels,
                      L->g1_channels, L->g0_mode, L->pe_first, L->pe_count,
                      L->lod_slot, F, origins, n_tiles, T, L->mip_dim[mip],
                      norm_lod, 0, 0, lvl, 0, 0.0f, 0.0f, e->g0[lvl],
                      e->g1[lvl], x);
  const float* w0 = e->dec;
  const float* w1 = w0 + (int64_t)H * F;
  const float* wout = w1 + (int64_t)H * H;
  if (e->calibrate)
  {
    o_mlp_forward_f32(F, H, O, L->activation, B, x, NULL, w0, w1, wout, mlp,
                      &e->max_a0, &e->max_a1);
    return;
  }
  o_draw_targets_f32(origins, n_tiles, T, L->mip_dim[mip], O, e->packages[mip], yt);
  o_mlp_forward_f32(F, H, O, L->activation, B, x, yt, w0, w1, wout, mlp, NULL, NULL);
  const float* yp = mlp + 4 * (int64_t)H * B;
  for (int o = 0; o < O; ++o)
  {
    double s = 0.0;
    for (int64_t t = 0; t < B; ++t)
    {
      const float r = yt[(int64_t)o * B + t] - yp[(int64_t)o * B + t];
      s += (double)r * (double)r;
    }
    e->sse += s;
    if (e->channel_sse) e->channel_sse[o] += s;
  }
  e->count += (double)B * O;
}

double o_eval_f32(const o_layout* L, float** g0, float** g1, const float* dec,
                  const uint8_t** packages, float* ws, double* channel_sse)
{
  o_eval_ctx e;
  memset(&e, 0, sizeof(e));
  e.g0 = g0;
  e.g1 = g1;
  e.dec = dec;
  e.packages = packages;
  e.ws = ws;
  e.channel_sse = channel_sse;
  o_walk_tiles(L, o_eval_batch, &e);
  return e.sse / e.count;
}

void o_calibrate_f32(const o_layout* L, float** g0, float** g1,
                     const float* dec, float* ws, float out[2])
{
  o_eval_ctx e;
  memset(&e, 0, sizeof(e));
  e.g0 = g0;
  e.g1 = g1;
  e.dec = dec;
  e.ws = ws;
  e.calibrate = 1;
  o_walk_tiles(L, o_eval_batch, &e);
  out[0] = e.max_a0;
  out[1] = e.max_a1;
}
