/* ntc_oracle.h - CPU restatement of the openntc training hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product (openntc_amd) never links or calls it.
 *
 * Every function states which reference lines it restates. Arithmetic is
 * written out operation by operation (explicit fmaf where an FMA is meant,
 * compiled with -ffp-contract=off) so that the GPU kernels, which follow the
 * same definitions, can be compared bit for bit where the operation order is
 * fixed (gather, targets, forward GEMMs, activations, loss, Adam, quantise,
 * pack, Lanczos) and within a stated tolerance where it is not (atomic
 * scatter order).
 *
 * Functions with suffix _f32 compute in float, _f64 in double (the FP64
 * oracle used for gradient checks and error statements).
 */
#ifndef NTC_ORACLE_H
#define NTC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mt19937 and the schedule distributions ---------------------------- */
typedef struct o_mt
{
  uint32_t s[624];
  int32_t i;
} o_mt;

void o_mt_seed(o_mt* m, uint32_t seed);
uint32_t o_mt_next(o_mt* m);
size_t o_mt_size(void);
/* [0,1) with 24 random bits */
float o_uniform01(o_mt* m);
/* (0,1] with 24 random bits */
float o_uniform01_open0(o_mt* m);
float o_uniform_range(o_mt* m, float lo, float hi);
int32_t o_uniform_int(o_mt* m, int32_t lo, int32_t hi);
int32_t o_bernoulli(o_mt* m, double p);

/* ---- layout (libopenntc.cpp:289-338, 881-891, 1209-1227) ---------------- */
typedef struct o_layout
{
  int32_t dim, mip_count, level_count, out_dim;
  int32_t mip_dim[12];
  int32_t g0_dim[6], g1_dim[6];
  int32_t g0_channels, g1_channels, g0_bits, g1_bits, g0_scale;
  int32_t feature_dim, feature_dim_padded, hidden_dim, out_dim_padded;
  int32_t g0_mode, pe_first, pe_count, lod_slot, activation;
  float g0_delta, g1_delta;
} o_layout;

/* profile 0..3 = Bpp_0_2, Bpp_0_5, Bpp_1_0, Bpp_2_25 */
int o_layout_profile(int32_t profile, int32_t dim, int32_t out_dim,
                     int32_t mip_min_dim, o_layout* out);
/* custom layout; feature_layout 0 reference, 1 bilinear */
int o_layout_custom(int32_t g0_bits, int32_t g1_bits, int32_t g0_ch,
                    int32_t g1_ch, int32_t g0_scale, int32_t hidden,
                    int32_t feature_layout, int32_t activation, int32_t dim,
                    int32_t out_dim, int32_t mip_min_dim, o_layout* out);
int32_t o_level_for_lod(const o_layout* L, int32_t lod);
/* payload bytes of the packed grids + decoder blob (Dump's blob sizes) */
uint64_t o_compressed_bytes(const o_layout* L);
uint64_t o_decoder_bytes(const o_layout* L);

/* ---- initialisation and schedule (libopenntc.cpp:471-504, 564-585,
 * 711-720) --------------------------------------------------------------- */
/* Fills grids level by level (g0 then g1) with U(lo, hi), then W0[H][F],
 * W1[H][H], Wout[16][H] with U(+-1/sqrt(fan_in)). g0/g1: arrays of
 * level_count pointers. Consumes draws from m. */
void o_init_params(o_mt* m, const o_layout* L, float lo, float hi,
                   float** g0, float** g1, float* w0, float* w1, float* wout);

typedef struct o_step
{
  int32_t lod, level, tile_dim, n_tiles;
  int32_t origins[16];
  float lr_grid, lr_decoder;
  float bias1, bias2, grid_bias1, grid_bias2;
  int32_t grid_update, lock;
} o_step;

/* Draws the schedule of steps [0, n_steps) from m; grid_counters holds the
 * per-level Adam step counters (level_count ints, start at 0). */
void o_schedule(o_mt* m, const o_layout* L, int32_t n_steps,
                int32_t grids_per_batch, float lr_grid_max,
                float lr_dec_max, o_step* out);
float o_cosine_annealing(float lr_min, float lr_max, int32_t t_max,
                         int32_t t_cur);

/* ---- Philox4x32-10 noise ------------------------------------------------- */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                     uint32_t out[4]);
/* Quantisation noise of grid (0 = g0, 1 = g1) element (cell, ch). */
float o_noise(uint32_t seed, uint32_t step, int32_t grid, int32_t level,
              uint32_t cell, int32_t ch, float delta);

/* ---- per-stage kernels, float and double --------------------------------- */
#define O_DECL(REAL, SFX)                                                      \
  void o_draw_features_##SFX(                                                  \
      int32_t g0_dim, int32_t g1_dim, int32_t c0, int32_t c1,                  \
      int32_t g0_mode, int32_t pe_first, int32_t pe_count, int32_t lod_slot,   \
      int32_t F, const int32_t* origins, int32_t n_tiles, int32_t T,           \
      int32_t M, float norm_lod, uint32_t seed, uint32_t step, int32_t level,  \
      int32_t noise_on, float d0, float d1, const REAL* g0, const REAL* g1,    \
      REAL* x);                                                                \
  void o_draw_targets_##SFX(const int32_t* origins, int32_t n_tiles,           \
                            int32_t T, int32_t M, int32_t O,                   \
                            const uint8_t* pkg, REAL* y);                      \
  void o_gemm_##SFX(int32_t op, int32_t n, int32_t m, int32_t k,               \
                    const REAL* a, const REAL* b, REAL* c);                    \
  void o_act_forward_##SFX(int32_t act, int64_t n, const REAL* z, REAL* a);    \
  void o_act_backward_##SFX(int32_t act, int64_t n, const REAL* z,             \
                            const REAL* g, REAL* out);                         \
  double o_loss_##SFX(int64_t n, REAL scale, const REAL* yt, const REAL* yp,   \
                      REAL* dy);                                               \
  void o_scatter_##SFX(int32_t g0_dim, int32_t g1_dim, int32_t c0, int32_t c1, \
                       int32_t g0_mode, const int32_t* origins,                \
                       int32_t n_tiles, int32_t T, int32_t M, const REAL* dx,  \
                       REAL* dg0, REAL* dg1);                                  \
  /* forward + loss + backward of one batch; returns SSE. Workspace holds   \
   * 6*H*B + 2*O*B + 2*F*B REALs. */                                          \
  double o_mlp_fwd_bwd_##SFX(int32_t F, int32_t H, int32_t O, int32_t act,     \
                             int64_t B, const REAL* x, const REAL* yt,         \
                             const REAL* w0, const REAL* w1,                   \
                             const REAL* wout, REAL* dw0, REAL* dw1,           \
                             REAL* dwout, REAL* dx, REAL* ws);                 \
  /* forward only, returns SSE; optional max|A0|, max|A1| (NULL to skip) */   \
  double o_mlp_forward_##SFX(int32_t F, int32_t H, int32_t O, int32_t act,     \
                             int64_t B, const REAL* x, const REAL* yt,         \
                             const REAL* w0, const REAL* w1,                   \
                             const REAL* wout, REAL* ws, REAL* max_a0,         \
                             REAL* max_a1);

O_DECL(float, f32)
O_DECL(double, f64)

/* ---- optimiser, grid maintenance, export (float) ------------------------- */
void o_adam_f32(int64_t n, float lr, float b1, float b2, float bias1,
                float bias2, const float* g, float* m, float* v, float* p);
void o_clamp_grid_f32(int64_t n, float delta, float* g);
void o_quantize_grid_f32(int64_t n, int32_t bits, float delta, float* g);
void o_quantize_pack_f32(int64_t n, int32_t bits, const float* in,
                         uint32_t* out);
void o_quantize_weights(const float* w, int32_t rows, int32_t cols,
                        float prescale, uint32_t* wq, float* ws);

/* ---- mips (ntc_kernel.cu:569-663) ---------------------------------------- */
void o_lanczos_weights(float w[12]);
void o_filter_lanczos(int32_t dim_src, int32_t c, const uint8_t* src,
                      uint8_t* dst);

/* ---- one full training step (libopenntc.cpp:568-790) ---------------------
 * Grids g0/g1 and their Adam state are per-level pointer arrays; the dense
 * grid update runs over the step's level. dec holds W0|W1|Wout (the decoder
 * layout of the C ABI), dec_m/dec_v its Adam state. packages: mip_count
 * pointers to [M][M][O] bytes. ws: o_step_workspace_elems(L, B) REALs.
 * Returns the step's SSE. */
size_t o_step_workspace_elems(const o_layout* L, int64_t max_batch);
double o_train_step_f32(const o_layout* L, const o_step* st, uint32_t seed,
                        uint32_t step_index, float** g0, float** g1,
                        float** g0m, float** g0v, float** g1m, float** g1v,
                        float* dec, float* dec_m, float* dec_v,
                        const uint8_t** packages, float* ws);
/* Lock step: quantise every level of both grids (libopenntc.cpp:601-619). */
void o_lock_grids_f32(const o_layout* L, float** g0, float** g1);
/* Full-coverage eval (libopenntc.cpp:972-1053): returns MSE, adds per-channel
 * SSE into channel_sse (O doubles) when not NULL. */
double o_eval_f32(const o_layout* L, float** g0, float** g1, const float* dec,
                  const uint8_t** packages, float* ws, double* channel_sse);
/* Calibration (libopenntc.cpp:898-970): out = {max|A0|, max|A1|}. */
void o_calibrate_f32(const o_layout* L, float** g0, float** g1,
                     const float* dec, float* ws, float out[2]);

/* ---- evaluation decode at reduced precision (ntc_oracle_decode.c) ------ */
float o_bf16_round(float x);
uint8_t o_e4m3_encode(float x);
float o_e4m3_decode(uint8_t c);
int32_t o_clamp_s8(float f);
/* s = {s_x, s_w0, s_w1, s_wout, s_a0, s_a1} */
void o_fp8_scales(int32_t F, int32_t H, int32_t O, const float* dec, float max_a0, float max_a1, float s[6]);
/* precision 0 FP32, 1 BF16, 2 FP8 (s8 = o_fp8_scales), 3 INT8 (blob = exported
 * decoder, rcp = {rcp_s_a1, rcp_s_a2}); x [F][B] FP32 features -> y [O][B]. */
void o_decode_f32(int32_t precision, int32_t act, int32_t F, int32_t Fp, int32_t H, int32_t O, int64_t B,
                  const float* dec, const float* s8, const uint8_t* blob, const float* rcp, const float* x,
                  float* y);
void o_decode_f64(int32_t act, int32_t F, int32_t H, int32_t O, int64_t B, const double* dec, const double* x,
                  double* y);

#ifdef __cplusplus
}
#endif

#endif
