/* ntc_hip.h - C ABI of the MI355X (gfx950) kernel library.
 *
 * This is the layer that replaces the reference's kernel launch functions
 * (libopenntc/ntc_kernel.cuh:1-78) and its three cuBLAS GEMM wrappers plus
 * the Saxpy/Sdot loss calls (libopenntc/libopenntc.cpp:397-417, 659-668).
 * Differences from those, by design:
 *   - every entry point returns an int status (0 = success, NTC_E_* or a
 *     positive hipError_t) instead of void;
 *   - every entry point takes an explicit stream;
 *   - quantisation noise is regenerated inside the gather from a counter-based
 *     generator (Philox4x32-10, Salmon et al., SC'11) keyed by
 *     (seed, step, grid, level, cell, channel) instead of being read from a
 *     per-step noise buffer filled by cuRAND;
 *   - buffers are caller-owned device memory; no entry point allocates except
 *     where stated.
 * All pointers are device pointers unless the comment says "host".
 * The header is plain C and needs no HIP header.
 */
#ifndef NTC_HIP_H
#define NTC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;

#define NTC_OK 0
#define NTC_E_INVALID (-1)    /* argument out of range / inconsistent shapes */
#define NTC_E_WORKSPACE (-2)  /* workspace too small */
#define NTC_E_NODEVICE (-3)   /* no usable gfx950 device */

/* g0 feature modes */
#define NTC_G0_CORNERS 0   /* 4*c0 slots: the 4 bilinear corners, unweighted */
#define NTC_G0_BILINEAR 1  /* c0 slots: bilinearly weighted */

/* activations */
#define NTC_ACT_HARDGELU 0
#define NTC_ACT_RELU 1

/* GEMM forms, row-major (libopenntc.cpp:395-417):
 *   AB : C[n][m] = A[n][k] * B[k][m]
 *   ATB: C[n][m] = A[k][n]^T * B[k][m]
 *   ABT: C[n][m] = A[n][k] * B[m][k]^T   (k split in chunks of
 *        NTC_GEMM_KCHUNK, each summed in order, chunk sums added in order) */
#define NTC_GEMM_AB 0
#define NTC_GEMM_ATB 1
#define NTC_GEMM_ABT 2
#define NTC_GEMM_KCHUNK 1024

typedef struct ntc_feature_desc
{
  int32_t g0_dim;          /* side of the active g0 level */
  int32_t g1_dim;          /* side of the active g1 level */
  int32_t g0_channels;
  int32_t g1_channels;
  int32_t g0_mode;         /* NTC_G0_CORNERS or NTC_G0_BILINEAR */
  int32_t pe_first_period; /* 8 for the reference layout */
  int32_t pe_count;        /* periods first, first/2, ...; 4 slots each */
  int32_t lod_slot;        /* 1: one slot holding tiles.norm_lod */
  int32_t feature_dim;     /* F; slots past the above are zero */
} ntc_feature_desc;

typedef struct ntc_tiles
{
  const int32_t* origins;  /* 2*n_tiles ints (x0, y0) per tile */
  int32_t n_tiles;
  int32_t tile_dim;        /* T: texels per tile side (wrapping at mip_dim) */
  int32_t mip_dim;         /* M: side of the sampled mip */
  float norm_lod;          /* lod / (mip_count - 1) */
} ntc_tiles;

typedef struct ntc_noise
{
  uint32_t seed;
  uint32_t step;      /* training step index */
  int32_t level;      /* active pyramid level */
  int32_t enabled;    /* 0: no noise (after the lock, eval) */
  float delta0;       /* g0 quantisation step 2/2^b0 */
  float delta1;       /* g1 quantisation step 2/2^b1 */
} ntc_noise;

/* ---- library ------------------------------------------------------------ */
const char* ntc_version(void);
/* Device memory helpers for FFI callers that have no HIP binding of their
 * own (test harnesses, ctypes). Synchronous; status as elsewhere. */
int ntc_dev_alloc(size_t bytes, void** out);
int ntc_dev_free(void* p);
int ntc_dev_upload(void* dst, const void* src, size_t bytes);
int ntc_dev_download(void* dst, const void* src, size_t bytes);
int ntc_dev_synchronize(void);
int ntc_dev_copy(void* dst, const void* src, size_t bytes); /* device to device */
/* 0 when a gfx950 device is usable; fills *count with the device count. */
int ntc_device_check(int32_t* count);
const char* ntc_status_string(int status);

/* ---- gather (replaces launch_draw_features, ntc_kernel.cu:75-214) -------- */
/* X[f][b*T*T + xy], f < F, B = n_tiles*T*T. */
int ntc_draw_features_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                          const ntc_noise* noise, const float* g0,
                          const float* g1, float* x, hipStream_t stream);
int ntc_draw_features_f64(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                          const ntc_noise* noise, const double* g0,
                          const double* g1, double* x, hipStream_t stream);

/* ---- targets (replaces launch_draw_targets, ntc_kernel.cu:216-259) ------- */
/* Y[o][b] = package[(y*M + x)*O + o] / 255 */
int ntc_draw_targets_f32(const ntc_tiles* tiles, int32_t out_dim,
                         const uint8_t* package, float* y, hipStream_t stream);
int ntc_draw_targets_f64(const ntc_tiles* tiles, int32_t out_dim,
                         const uint8_t* package, double* y, hipStream_t stream);

/* ---- GEMM (replaces matmulAB/ATB/ABT + cublasSgemm) ----------------------- */
size_t ntc_gemm_workspace_bytes(int32_t op, int32_t n, int32_t m, int32_t k,
                                int32_t elem_size);
int ntc_gemm_f32(int32_t op, int32_t n, int32_t m, int32_t k, const float* a,
                 const float* b, float* c, void* workspace,
                 size_t workspace_bytes, hipStream_t stream);
int ntc_gemm_f64(int32_t op, int32_t n, int32_t m, int32_t k, const double* a,
                 const double* b, double* c, void* workspace,
                 size_t workspace_bytes, hipStream_t stream);

/* ---- activations (replace forward/backward_hardgelu, ntc_kernel.cu:13-48) */
int ntc_act_forward_f32(int32_t act, int64_t n, const float* z, float* a,
                        hipStream_t stream);
int ntc_act_backward_f32(int32_t act, int64_t n, const float* z,
                         const float* g, float* out, hipStream_t stream);
int ntc_act_forward_f64(int32_t act, int64_t n, const double* z, double* a,
                        hipStream_t stream);
int ntc_act_backward_f64(int32_t act, int64_t n, const double* z,
                         const double* g, double* out, hipStream_t stream);

/* ---- loss (replaces Saxpy + Sdot + scalar_product, libopenntc.cpp:659-668)
 * r = y_target - y_pred; dy = scale * r; *sse = sum r^2 accumulated in FP64
 * (two-pass, deterministic). workspace: ntc_loss_workspace_bytes(n). */
size_t ntc_loss_workspace_bytes(int64_t n);
int ntc_loss_f32(int64_t n, float scale, const float* y_target,
                 const float* y_pred, float* dy, double* sse, void* workspace,
                 hipStream_t stream);
int ntc_loss_f64(int64_t n, double scale, const double* y_target,
                 const double* y_pred, double* dy, double* sse,
                 void* workspace, hipStream_t stream);

/* Per-channel SSE: sse[o] += sum_b (y_target[o][b] - y_pred[o][b])^2 in FP64,
 * O <= 16, deterministic. workspace: ntc_channel_sse_workspace_bytes(). */
size_t ntc_channel_sse_workspace_bytes(void);
int ntc_channel_sse_f32(int32_t out_dim, int64_t batch, const float* y_target,
                        const float* y_pred, double* sse, void* workspace,
                        hipStream_t stream);

/* ---- latent-gradient scatter (replaces accumulate_grid_gradients,
 * ntc_kernel.cu:261-373). Adds into dg0/dg1; the caller zeroes them. -------- */
int ntc_accumulate_grid_gradients_f32(const ntc_feature_desc* fd,
                                      const ntc_tiles* tiles, const float* dx,
                                      float* dg0, float* dg1,
                                      hipStream_t stream);
int ntc_accumulate_grid_gradients_f64(const ntc_feature_desc* fd,
                                      const ntc_tiles* tiles, const double* dx,
                                      double* dg0, double* dg1,
                                      hipStream_t stream);

/* ---- optimiser and grid maintenance ------------------------------------- */
/* Dense Adam (update_adam, ntc_kernel.cu:375-410). */
int ntc_adam_f32(int64_t n, float lr, float beta1, float beta2, float bias1,
                 float bias2, const float* grad, float* m, float* v, float* p,
                 hipStream_t stream);
/* Fused grid update: Adam, then clamp p to [lo, hi], then grad = 0; one pass
 * over p, m, v, grad (update_adam + clamp_grid + the next step's FillZero). */
int ntc_grid_update_f32(int64_t n, float lr, float beta1, float beta2,
                        float bias1, float bias2, float lo, float hi,
                        float* grad, float* m, float* v, float* p,
                        hipStream_t stream);
/* The same for both grids of a pyramid level in one launch (sizes multiples
 * of 4, 16-byte aligned buffers); per-grid clamp bounds. */
int ntc_grid_update2_f32(int64_t n0, float* grad0, float* m0, float* v0,
                         float* p0, float lo0, float hi0, int64_t n1,
                         float* grad1, float* m1, float* v1, float* p1,
                         float lo1, float hi1, float lr, float beta1,
                         float beta2, float bias1, float bias2,
                         hipStream_t stream);
/* clamp_grid (ntc_kernel.cu:468-482): g = clamp(g, -1, 1 - delta) */
int ntc_clamp_grid_f32(int64_t n, float delta, float* g, hipStream_t stream);
/* quantize_grid (ntc_kernel.cu:448-466) */
int ntc_quantize_grid_f32(int64_t n, int32_t bits, float delta, float* g,
                          hipStream_t stream);
/* quantize_pack (ntc_kernel.cu:486-507): n*bits must be a multiple of 32 */
int ntc_quantize_pack_f32(int64_t n, int32_t bits, const float* in,
                          uint32_t* out, hipStream_t stream);
/* max_abs (ntc_kernel.cu:511-531): *result = max(*result, max|x|) */
int ntc_max_abs_f32(int64_t n, const float* x, float* result,
                    hipStream_t stream);
/* Adam state and constants of the active level's two grids (update_adam +
 * clamp_grid, ntc_kernel.cu:375-410 and 468-482). */
typedef struct ntc_grid_adam
{
  float* m0; /* g0 moments */
  float* v0;
  float* m1; /* g1 moments */
  float* v1;
  float lr, beta1, beta2, bias1, bias2;
  float lo0, hi0; /* clamp ranges of g0 and g1 */
  float lo1, hi1;
} ntc_grid_adam;

/* The dense Adam + clamp + gradient reset of ntc_grid_update2_f32 over the
 * whole level of g0 (fd->g0_dim^2 x g0_channels) and g1, for gradients that
 * are zero outside the tile footprints (the latent scatter of one step):
 * cells outside every tile's footprint rectangle take the zero-gradient step
 * without reading or writing dg, 24 instead of 32 bytes per parameter.
 * Same result as ntc_grid_update2_f32, bit for bit. Channel counts must be
 * multiples of 4 and n_tiles <= 16. */
int ntc_grid_update_tiles_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                              float* g0, float* dg0, float* g1, float* dg1,
                              const ntc_grid_adam* adam, hipStream_t stream);
/* One part of ntc_grid_update_tiles_f32 (together the two parts give its
 * result bit for bit): NTC_TILES_COLD, the cells outside every footprint
 * rectangle (dg not touched; this part reads and writes nothing that the
 * same step's training kernel reads or writes, so it may run on another
 * stream beside it), or NTC_TILES_HOT, the cells inside (dg read and
 * cleared). blocks: launch width, <= 0 for the default. Grid sides must be
 * powers of two and channel counts at most 16. No counterpart in the
 * reference, whose update is one dense pass (ntc_kernel.cu:375-410). */
#define NTC_TILES_COLD 1
#define NTC_TILES_HOT 2
int ntc_grid_update_tiles_part_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                                   float* g0, float* dg0, float* g1, float* dg1,
                                   const ntc_grid_adam* adam, int32_t part,
                                   int32_t blocks, hipStream_t stream);
/* The tiles update with the zero-gradient steps of cells outside the
 * footprints deferred (no counterpart in the reference; its update is one
 * dense pass, ntc_kernel.cu:375-410, whose results these reproduce bit for
 * bit). last0/last1 (device, one int per g0 / g1 cell of the level) count the
 * level's Adam steps applied to each cell; steps (device) holds
 * {lr, bias1, bias2} of the level's step s at [3 s]; adam supplies m, v,
 * beta1, beta2 and the clamp bounds (and, for the hot part, this step's lr
 * and bias factors).
 * Catch-up: applies the zero-gradient steps last .. upto - 1 (in order) to
 * the cells of the footprint rectangles of tiles, leaving out those in the
 * rectangles of skip (null: none; the previous step's tiles on the same
 * level, whose hot part brings those cells current, so that this catch-up
 * can run beside it), or, with tiles null, to every cell of both grids (last
 * is not changed: after an every-cell catch-up set it to upto with
 * ntc_fill_i32).
 * Hot: the level's step `step` on the cells of the rectangles (gradient read
 * and cleared, as NTC_TILES_HOT), which must be current (caught up to step),
 * then last = step + 1 there. Grid sides powers of two, channel counts
 * multiples of 4 up to 16, n_tiles <= 16. */
int ntc_grid_lazy_catchup_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                              const ntc_tiles* skip, float* g0, float* g1,
                              const ntc_grid_adam* adam,
                              const float* steps, int32_t upto,
                              const int32_t* last0, const int32_t* last1,
                              hipStream_t stream);
int ntc_grid_lazy_hot_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                          float* g0, float* dg0, float* g1, float* dg1,
                          const ntc_grid_adam* adam, int32_t step, int32_t* last0,
                          int32_t* last1, hipStream_t stream);
/* ntc_grid_lazy_hot_f32 of one step and ntc_grid_lazy_catchup_f32 of the next
 * (next_* arguments; skip: the first step's tiles when both are on one level,
 * so that the two touch different cells) in one launch, on the blocks of each
 * part: the step's tail on one stream, with no second stream to wait for.
 * Not across a flush (as the catch-up issued ahead). Context::Train keeps
 * the two-stream form, which overlaps the catch-up with the decoder update
 * as well (0.477 against 0.488 ms per step at 4096^2). No counterpart in the
 * reference. */
int ntc_grid_lazy_hot_catchup_f32(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                                  float* g0, float* dg0, float* g1, float* dg1,
                                  const ntc_grid_adam* adam, int32_t step,
                                  int32_t* last0, int32_t* last1,
                                  const ntc_feature_desc* next_fd,
                                  const ntc_tiles* next_tiles, const ntc_tiles* skip,
                                  float* next_g0, float* next_g1,
                                  const ntc_grid_adam* next_adam, const float* steps,
                                  int32_t upto, int32_t* next_last0,
                                  int32_t* next_last1, hipStream_t stream);
/* out[i] = value, i < n (kernel fill) */
int ntc_fill_i32(int64_t n, int32_t value, int32_t* out, hipStream_t stream);
/* zero-fill by kernel (graph-safe replacement for memset) */
int ntc_fill_zero(void* p, size_t bytes, hipStream_t stream);

/* ---- fused BF16 training step ------------------------------------------------
 * One persistent launch per step does the gather, the decoder forward, the
 * loss, the hand-written backward and the latent-gradient scatter
 * (replacing launch_draw_features .. launch_accumulate_grid_gradients and the
 * nine SGEMMs, libopenntc.cpp:621-709); ntc_step_forward / ntc_step_backward
 * below are the same step as a forward / backward pair. MFMA operands are BF16, accumulation FP32,
 * master latents and weights stay FP32. */
typedef struct ntc_fused_args
{
  ntc_feature_desc feat;
  ntc_tiles tiles;
  ntc_noise noise;
  int32_t hidden_dim, out_dim, activation;
  const float* g0;          /* active level, FP32 masters */
  const float* g1;
  const uint16_t* weights;  /* BF16 tables from ntc_pack_decoder_bf16 */
  const uint8_t* targets;   /* target package of the sampled mip [M][M][O] */
  float loss_scale;         /* -2 / (O * B) */
  int32_t scatter;          /* 1: add latent gradients into dg0/dg1 */
  float* dg0;
  float* dg1;
  float* dw_partials;       /* [blocks][decoder elems] */
  double* sse_partials;     /* [blocks] */
  int32_t blocks;           /* == ntc_fused_blocks_for(...) */
} ntc_fused_args;

/* 1 when a fused kernel is compiled for this feature layout / width. */
int ntc_fused_supported(const ntc_feature_desc* fd, int32_t hidden_dim,
                        int32_t activation);
size_t ntc_fused_table_elems(const ntc_feature_desc* fd, int32_t hidden_dim);
int ntc_fused_max_blocks(int32_t* blocks);
int ntc_fused_blocks_for(const ntc_feature_desc* fd, int32_t hidden_dim,
                         int32_t activation, int64_t batch,
                         int32_t max_blocks);
/* FP32 decoder W0[H][F] | W1[H][H] | Wout[16][H] -> BF16 MFMA tables. */
int ntc_pack_decoder_bf16(const ntc_feature_desc* fd, int32_t hidden_dim,
                          int32_t out_dim, const float* dec, uint16_t* table,
                          hipStream_t stream);
/* tiles.tile_dim must be a power of two (NTC_E_INVALID otherwise), as every
 * mip side of a Context is. */
int ntc_fused_train_step(const ntc_fused_args* args, hipStream_t stream);
/* The same step as a forward / backward pair, split where the reference
 * splits it, at the loss (forward: draw_features + 3 GEMMs + activations,
 * libopenntc.cpp:621-655; backward: 6 GEMMs + activation backward +
 * accumulate_grid_gradients, :672-709 and ntc_kernel.cu:261-373). Arguments
 * as ntc_fused_train_step's.
 * ntc_step_forward writes yhat [O][B] (FP32, columns in batch order: tile,
 * then rows of the tile); with targets it also writes sse_partials (the SSE
 * of the batch is their ordered sum); g0/g1 are only read.
 * ntc_step_backward takes dyhat = dL/dYhat [O][B] in place of the loss
 * gradient the training step forms from the targets (loss_scale and targets
 * are not used), and writes the decoder-gradient partials and (scatter = 1)
 * adds the latent gradients into dg0/dg1. Nothing is kept between the two
 * calls: the backward gathers and runs the forward again (recompute policy).
 * Given dyhat = loss_scale * (Y* - yhat) formed in FP32 from the forward's
 * yhat, the pair gives the training step's results. */
int ntc_step_forward(const ntc_fused_args* args, float* yhat, hipStream_t stream);
int ntc_step_backward(const ntc_fused_args* args, const float* dyhat, hipStream_t stream);
/* Ordered sum of the workgroup partials into grad [H*F + H*H + 16*H] (the
 * decoder update's gradient, bit for bit, without the Adam step); sse_out
 * (may be NULL) receives the ordered sum of sse_partials. */
int ntc_reduce_decoder_partials(int32_t hidden_dim, int32_t feature_dim,
                                int32_t blocks, const float* dw_partials,
                                float* grad, const double* sse_partials,
                                double* sse_out, hipStream_t stream);
/* ntc_fused_train_step followed by ntc_reduce_decoder_partials, with no
 * workspace argument and no allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under graph capture,
 * serving loops, harnesses that forbid hipMalloc). args->dw_partials and
 * args->sse_partials must be NULL; the partials live in a device array of
 * the library, one per GPU, held until the process ends, of
 * ntc_fused_train_step_noalloc_max() floats (256 x 32768): blocks * (decoder
 * elements) above that, or blocks above 256, return NTC_E_WORKSPACE. Writes
 * grad (the decoder gradient, as ntc_reduce_decoder_partials) and *sse_out
 * (may be NULL); latent gradients as ntc_fused_train_step. Calls that
 * overlap on one GPU share the array and race. */
size_t ntc_fused_train_step_noalloc_max(void);
int ntc_fused_train_step_noalloc(const ntc_fused_args* args, float* grad,
                                 double* sse_out, hipStream_t stream);
/* The training step of ntc_fused_train_step in three launches (replaces the
 * same reference stages: draw_features, the forward and backward GEMMs,
 * activations, loss and accumulate_grid_gradients, libopenntc.cpp:621-709,
 * ntc_kernel.cu:75-373): ntc_split_gather writes the features of every texel
 * (with their quantisation noise) into BF16 rows of the workspace,
 * ntc_split_mlp runs the decoder forward, loss and backward on BF16 MFMA
 * (decoder-gradient and SSE partials as ntc_fused_train_step's) and writes
 * the dX of the grid slots into FP32 rows of the workspace, ntc_split_scatter
 * adds the latent gradients from those rows (scatter = 1; nothing
 * otherwise). ntc_split_train_step issues the three in order on the stream;
 * a caller may issue them one by one (with its own work or events between
 * them) as long as they run in that order on one step's arguments.
 * Compiled for the H = 128 ReLU layouts (ntc_split_supported). Arguments as
 * ntc_fused_train_step's, except that blocks must equal
 * ntc_split_blocks_for(...) (the MLP launch's workgroups, one partial slot
 * each) and the workspace (256-byte aligned, at least
 * ntc_split_workspace_bytes(...) bytes) holds the feature rows first, then
 * the dX rows, both in pass order (128-texel passes). */
int ntc_split_supported(const ntc_feature_desc* fd, int32_t hidden_dim,
                        int32_t activation);
int ntc_split_blocks_for(const ntc_feature_desc* fd, int32_t hidden_dim,
                         int32_t activation, int64_t batch,
                         int32_t max_blocks);
size_t ntc_split_workspace_bytes(const ntc_feature_desc* fd, int32_t hidden_dim,
                                 int32_t activation, int64_t batch);
int ntc_split_train_step(const ntc_fused_args* args, void* workspace,
                         size_t workspace_bytes, hipStream_t stream);
int ntc_split_gather(const ntc_fused_args* args, void* workspace,
                     size_t workspace_bytes, hipStream_t stream);
int ntc_split_mlp(const ntc_fused_args* args, void* workspace,
                  size_t workspace_bytes, hipStream_t stream);
int ntc_split_scatter(const ntc_fused_args* args, void* workspace,
                      size_t workspace_bytes, hipStream_t stream);
/* What follows the MLP launch of one training step, in one launch on one
 * stream: ntc_split_scatter's latent gradients (args->scatter = 1), the
 * catch-up of ntc_grid_lazy_catchup_f32 for the next step's tiles
 * (catch_tiles non-null and catch_upto > 0; catch_skip as its skip, which
 * must hold this step's tiles when the next step is on the same level) and
 * ntc_decoder_update_bf16 of this step's partials (dec_p non-null; blocks,
 * dw_partials and sse_partials from args, the table layout from dec_feat).
 * The same arithmetic as those calls (the scatter's FP32 atomics in any
 * order, as there); the three touch disjoint memory, so their workgroups run
 * beside each other with no event or second stream between them. Replaces
 * the reference's accumulate_grid_gradients, zero-gradient update_adam /
 * clamp_grid of the level and the decoder's update_adam
 * (libopenntc.cpp:697-778, ntc_kernel.cu:261-410, 468-482). */
typedef struct ntc_split_tail_args
{
  const ntc_feature_desc* catch_feat;
  const ntc_tiles* catch_tiles;
  const ntc_tiles* catch_skip;
  float* catch_g0;
  float* catch_g1;
  const ntc_grid_adam* catch_adam;
  const float* catch_steps;
  int32_t catch_upto;
  int32_t* catch_last0;
  int32_t* catch_last1;
  const ntc_feature_desc* dec_feat;
  float dec_lr, dec_beta1, dec_beta2, dec_bias1, dec_bias2;
  float* dec_grad;
  float* dec_m;
  float* dec_v;
  float* dec_p;
  uint16_t* dec_table;
  double* sse_out;
} ntc_split_tail_args;
int ntc_split_tail(const ntc_fused_args* args, void* workspace, size_t workspace_bytes,
                   const ntc_split_tail_args* tail, hipStream_t stream);
/* Sums the workgroup partials in order into grad, then Adam on W0 | W1 |
 * Wout rows < O (update_adam x3, libopenntc.cpp:722-754); sse_out (may be
 * NULL) receives the ordered sum of sse_partials. */
int ntc_decoder_update(int32_t hidden_dim, int32_t feature_dim,
                       int32_t out_dim, int32_t blocks,
                       const float* dw_partials, float lr, float beta1,
                       float beta2, float bias1, float bias2, float* grad,
                       float* m, float* v, float* p,
                       const double* sse_partials, double* sse_out,
                       hipStream_t stream);
/* ntc_decoder_update that also brings the BF16 tables of
 * ntc_pack_decoder_bf16 (packed from p before, same fd / hidden_dim / out_dim)
 * up to date with the updated parameters: the table refresh of the fused step
 * without a separate pass over the tables. Bit-identical to
 * ntc_decoder_update followed by ntc_pack_decoder_bf16. */
int ntc_decoder_update_bf16(const ntc_feature_desc* fd, int32_t hidden_dim,
                            int32_t out_dim, int32_t blocks,
                            const float* dw_partials, float lr, float beta1,
                            float beta2, float bias1, float bias2, float* grad,
                            float* m, float* v, float* p, uint16_t* table,
                            const double* sse_partials, double* sse_out,
                            hipStream_t stream);

/* ---- several independent training sets in one launch per stage -------------
 * No counterpart in the reference (one Context trains one texture set,
 * libopenntc.cpp:536-879); these batch the fused step above over sets of one
 * decoder shape. Per-step parameters live in device tables indexed by a
 * device-side step counter (*step), so a captured graph of one step replays
 * every step unchanged; ntc_step_advance increments the counter. */
typedef struct ntc_dec_set
{
  const float* dw_partials;   /* [blocks][decoder elems] of this set */
  float* grad;
  float* m;
  float* v;
  float* p;
  const double* sse_partials; /* [blocks] */
  double* sse_out;            /* may be NULL */
} ntc_dec_set;

typedef struct ntc_dec_step
{
  int32_t blocks;             /* workgroups the set had this step */
  float lr, beta1, beta2, bias1, bias2;
} ntc_dec_step;

typedef struct ntc_pack_set
{
  const float* dec;
  uint16_t* table;
} ntc_pack_set;

/* Size of one opaque per-(step, set) record of the fused launch, and the
 * record of the arguments a (a->blocks is not used: the set's workgroup
 * range comes from wg_first; tile_dim a power of two as above). */
size_t ntc_fused_params_bytes(void);
int ntc_fused_params_pack(const ntc_fused_args* a, void* record);
/* Texels per workgroup pass of the compiled multi-set kernels of a shape
 * (sizes[0..n), n <= 2 returned; 0: none). */
int ntc_fused_pass_sizes(const ntc_feature_desc* fd, int32_t hidden_dim,
                         int32_t activation, int32_t* sizes);
/* params: device records [steps][n_sets]; wg_first: [steps][n_sets + 1]
 * workgroup offsets, set s of step t taking workgroups
 * [wg_first[t][s], wg_first[t][s + 1]) (its partial slots 0..count-1);
 * grid >= wg_first[t][n_sets] for every step. */
int ntc_fused_train_sets(const ntc_feature_desc* fd, int32_t hidden_dim,
                         int32_t activation, int32_t pass_texels,
                         const void* params, const int32_t* wg_first,
                         int32_t n_sets, const int32_t* step, int32_t grid,
                         hipStream_t stream);
/* ntc_decoder_update for every set: sets[n_sets] fixed, steps[steps][n_sets]. */
int ntc_decoder_update_sets(int32_t hidden_dim, int32_t feature_dim,
                            int32_t out_dim, const ntc_dec_set* sets,
                            const ntc_dec_step* steps, int32_t n_sets,
                            const int32_t* step, hipStream_t stream);
/* ntc_pack_decoder_bf16 for every set (sets: device array). */
int ntc_pack_decoder_bf16_sets(const ntc_feature_desc* fd, int32_t hidden_dim,
                               int32_t out_dim, const ntc_pack_set* sets,
                               int32_t n_sets, hipStream_t stream);
/* ntc_grid_update_tiles_f32 for every set: records [steps][n_sets] from
 * ntc_grid_update_tiles_pack (items: the float4 count of the record, for
 * sizing block ranges), blk_first [steps][n_sets + 1] block offsets (a set
 * without a grid update that step gets an empty range). */
size_t ntc_grid_update_tiles_bytes(void);
int ntc_grid_update_tiles_pack(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                               float* g0, float* dg0, float* g1, float* dg1,
                               const ntc_grid_adam* adam, void* record,
                               int64_t* items);
int ntc_grid_update_tiles_sets(const void* records, const int32_t* blk_first,
                               int32_t n_sets, const int32_t* step,
                               int32_t grid, hipStream_t stream);
int ntc_step_advance(int32_t* step, hipStream_t stream);
/* The deferred form of the grid update for several sets (no counterpart in
 * the reference; the multi-set form of ntc_grid_lazy_catchup_f32 and
 * ntc_grid_lazy_hot_f32): records of ntc_grid_lazy_bytes() from
 * ntc_grid_lazy_pack (the step's tiles, its Adam constants; upto: for the
 * catch-up the number of the level's steps before this one, for the gradient
 * step its own index; skip: tiles whose rectangles the catch-up leaves out,
 * as for ntc_grid_lazy_catchup_f32, or NULL; rect_blocks: the blocks per
 * footprint rectangle the single-set launch would use), [steps][n_sets], and
 * blk_first as for ntc_grid_update_tiles_sets, split over each set's
 * rectangle blocks; slots: twice the most tiles of a record (rectangle
 * slots, at most 32). A step's cells are caught up before its
 * ntc_fused_train_sets reads them (for example beside the previous step's
 * gradient step, leaving out its footprint), the gradient step runs after
 * the fused launch. */
size_t ntc_grid_lazy_bytes(void);
int ntc_grid_lazy_pack(const ntc_feature_desc* fd, const ntc_tiles* tiles,
                       const ntc_tiles* skip, float* g0, float* dg0, float* g1,
                       float* dg1, const ntc_grid_adam* adam, const float* steps,
                       int32_t upto, int32_t* last0, int32_t* last1, void* record,
                       int32_t* rect_blocks);
int ntc_grid_lazy_catchup_sets(const void* records, const int32_t* blk_first,
                               int32_t n_sets, const int32_t* step, int32_t grid,
                               int32_t slots, hipStream_t stream);
int ntc_grid_lazy_hot_sets(const void* records, const int32_t* blk_first,
                           int32_t n_sets, const int32_t* step, int32_t grid,
                           int32_t slots, hipStream_t stream);

/* ---- evaluation decode at reduced precision (Eval, libopenntc.cpp:972-1053,
 * at the precisions of EvalEx; INT8 = the deployed HLSL decoder,
 * openntc-gui/ntc_decode.hlsli:112-115, 400-502) ----------------------------
 * A decode table holds the weights of one precision (built on the host);
 * ntc_decode_sse_f32 decodes FP32 features X [F][B] from ntc_draw_features_f32
 * and adds each channel's squared error against Y* [O][B] to sse[16] (FP64,
 * fixed order). FP32: one FMA chain per output in ascending order (the
 * plain forward). BF16: weights, features and hidden activations rounded to
 * BF16, FP32 accumulation. FP8: OCP E4M3 operands, per-layer scales
 * (features 1/256, weights max|W|/448, hidden activations max|a|/448 from
 * max_a0/max_a1), exact products, FP32 accumulation. INT8: int8 features
 * clamp(round(128 x)), the exported blob's int8 weights and row scales,
 * int32 dot products, per-layer requantisation with rcp_s_a1/rcp_s_a2.
 * BF16_MFMA / FP8_MFMA: BF16's / FP8's rounding points on the matrix cores
 * (v_mfma_f32_16x16x32_bf16 / _fp8_fp8, FP32 accumulation in the MFMA's
 * order, so the outputs differ from the fixed-order chains by accumulation
 * rounding); their tables hold BF16 / E4M3 weights; feature_dim_padded unused. The FP32, BF16,
 * FP8 and INT8 forms are reproducible bit for bit on the CPU. */
#define NTC_PREC_FP32 0
#define NTC_PREC_BF16 1
#define NTC_PREC_FP8 2
#define NTC_PREC_INT8 3
#define NTC_PREC_BF16_MFMA 4
#define NTC_PREC_FP8_MFMA 5
size_t ntc_decode_table_bytes(int32_t precision, int32_t feature_dim, int32_t feature_dim_padded,
                              int32_t hidden_dim);
/* FP32 / BF16 / FP8 table (host memory) from the FP32 decoder
 * W0[H][F] | W1[H][H] | Wout[16][H] (host memory); H is 64 or 128. */
int ntc_build_decode_table_f(int32_t precision, int32_t feature_dim, int32_t hidden_dim, int32_t out_dim,
                             const float* dec, float max_a0, float max_a1, void* table);
/* INT8 table (host memory) from the exported decoder blob
 * W0 i8 [H][Fp] | W1 i8 [H][H] | Wout i8 [16][H] | f32 scales W0s[H] W1s[H] Wouts[16]. */
int ntc_build_decode_table_int8(int32_t feature_dim, int32_t feature_dim_padded, int32_t hidden_dim,
                                int32_t out_dim, const uint8_t* blob, float rcp_s_a1, float rcp_s_a2,
                                void* table);
size_t ntc_decode_workspace_bytes(int64_t batch);
/* table: device copy; y_out [O][B] may be NULL. */
int ntc_decode_sse_f32(int32_t precision, int32_t activation, int32_t feature_dim, int32_t feature_dim_padded,
                       int32_t hidden_dim, int32_t out_dim, int64_t batch, const void* table, const float* x,
                       const float* y_target, double* sse, float* y_out, void* workspace, hipStream_t stream);
/* The same for one batch of tiles whose features and targets the decode
 * kernel forms itself (the gather of ntc_draw_features_f32 with noise off and
 * the targets of ntc_draw_targets_f32, bit for bit), so they never go through
 * device memory: g0/g1 the level's grids, package the mip's targets
 * [M][M][O]; workspace: ntc_decode_workspace_bytes(batch). */
int ntc_decode_tiles_sse_f32(int32_t precision, int32_t activation, const ntc_feature_desc* fd,
                             const ntc_tiles* tiles, const float* g0, const float* g1,
                             const uint8_t* package, int32_t feature_dim_padded, int32_t hidden_dim,
                             int32_t out_dim, const void* table, double* sse, void* workspace,
                             hipStream_t stream);

/* ---- mip targets (prepare_tex / filter_lanczos / package_tex,
 * ntc_kernel.cu:533-686) ---------------------------------------------------- */
/* out[y][x][i] = tex[y][x][cmap[i]] for i < count of leading valid cmap
 * entries (cmap is host memory, 4 ints, -1 terminates). */
int ntc_prepare_tex(int32_t dim, int32_t src_channels, const int32_t* cmap,
                    const uint8_t* tex, uint8_t* out, hipStream_t stream);
/* Separable 2:1 Lanczos-3 downsample of a dim_src^2 x channels uint8 image
 * (channels <= 4); tmp holds (dim_src/2)*dim_src*channels floats. */
int ntc_filter_lanczos(int32_t dim_src, int32_t channels, const uint8_t* src,
                       float* tmp, uint8_t* dst, hipStream_t stream);
/* Interleave sources into [dim][dim][out_channels]. sources, source_channels,
 * feat_source and feat_channel are host arrays. */
int ntc_package_tex(int32_t dim, int32_t out_channels, int32_t n_sources,
                    const uint8_t* const* sources,
                    const int32_t* source_channels,
                    const int32_t* feat_source, const int32_t* feat_channel,
                    uint8_t* out, hipStream_t stream);

/* ---- Context shim --------------------------------------------------------
 * Flat C entry points over openntc::Context (include/libopenntc/libopenntc.h)
 * for callers that bind through an FFI (ctypes, cgo, JNI). Status values are
 * openntc::Result cast to int. */
typedef struct ntc_context ntc_context;

typedef struct ntc_layout_info
{
  int32_t g0_bits, g1_bits, g0_channels, g1_channels, g0_scale, hidden_dim;
  int32_t feature_layout; /* 0 reference, 1 bilinear */
  int32_t activation;     /* NTC_ACT_* */
} ntc_layout_info;

typedef struct ntc_init_ex
{
  int32_t profile;
  int32_t use_layout;
  ntc_layout_info layout;
  int32_t min_dimension;
  int32_t mip_min_dim;
  uint32_t seed;
} ntc_init_ex;

typedef struct ntc_train_ex
{
  int32_t steps, grids_per_batch;
  float lr_grid, lr_decoder, init_lo, init_hi;
  int32_t path; /* 0 fused bf16, 1 plain fp32 */
} ntc_train_ex;

typedef struct ntc_progress
{
  int32_t phase, result, batches_complete, total_batches;
} ntc_progress;

/* Derived sizes of a loaded context. */
typedef struct ntc_ctx_layout
{
  int32_t dim, mip_count, level_count, out_dim;
  int32_t mip_dim[12];
  int32_t g0_dim[6], g1_dim[6];
  int32_t g0_channels, g1_channels, g0_bits, g1_bits, g0_scale;
  int32_t feature_dim, feature_dim_padded, hidden_dim, out_dim_padded;
  int32_t activation;
  ntc_feature_desc feat; /* with g0_dim/g1_dim of level 0 */
  float g0_delta, g1_delta;
  int32_t lock_step, total_steps, grids_per_batch, batch_index;
  uint32_t seed;
} ntc_ctx_layout;

/* One step of the precomputed schedule. */
typedef struct ntc_step_info
{
  int32_t lod, level, tile_dim, n_tiles;
  int32_t origins[16];
  float lr_grid, lr_decoder;
  float bias1, bias2, grid_bias1, grid_bias2;
  int32_t grid_update; /* 1 before the lock */
  int32_t lock;        /* 1 on the lock step */
} ntc_step_info;

ntc_context* ntc_ctx_create(void);
void ntc_ctx_destroy(ntc_context* ctx);
int ntc_ctx_init(ntc_context* ctx, int32_t profile);
int ntc_ctx_init_ex(ntc_context* ctx, const ntc_init_ex* info);
/* rgba8: host pointers, count entries; channels: strings like "RGB". */
int ntc_ctx_load_textures(ntc_context* ctx, int32_t count,
                          const uint8_t* const* rgba8,
                          const int32_t* semantics,
                          const char* const* channels, int32_t dim);
int ntc_ctx_load_manifest(ntc_context* ctx, const char* path);
void ntc_ctx_unload_manifest(ntc_context* ctx);
int ntc_ctx_begin_training(ntc_context* ctx, int32_t quality,
                           int32_t grids_per_batch);
int ntc_ctx_begin_training_ex(ntc_context* ctx, const ntc_train_ex* info);
int ntc_ctx_train(ntc_context* ctx, int32_t n, ntc_progress* out);
int ntc_ctx_train_until_complete(ntc_context* ctx, ntc_progress* out);
int ntc_ctx_eval(ntc_context* ctx, int32_t precision, double* mse,
                 double* psnr, double* channel_sse /* host, 16 or NULL */);
int ntc_ctx_calibrate(ntc_context* ctx, float headroom, float* out4);
int ntc_ctx_get_last_error(ntc_context* ctx);
int ntc_ctx_get_layout(ntc_context* ctx, ntc_ctx_layout* out);
int ntc_ctx_step_info(ntc_context* ctx, int32_t step, ntc_step_info* out);
int ntc_ctx_feature_level_for_lod(ntc_context* ctx, int32_t lod);
/* which: 0 g0, 1 g1, 2 g0 Adam m, 3 g0 Adam v, 4 g1 m, 5 g1 v, 6 dg0, 7 dg1 */
int ntc_ctx_read_grid(ntc_context* ctx, int32_t which, int32_t level,
                      float* host);
int ntc_ctx_write_grid(ntc_context* ctx, int32_t which, int32_t level,
                       const float* host);
/* which: 0 weights, 1 Adam m, 2 Adam v; layout W0[H][F] | W1[H][H] |
 * Wout[16][H] */
int ntc_ctx_read_decoder(ntc_context* ctx, int32_t which, float* host);
int ntc_ctx_write_decoder(ntc_context* ctx, int32_t which, const float* host);
int ntc_ctx_read_package(ntc_context* ctx, int32_t mip, uint8_t* host);
/* Copies the compressed view: packed g0/g1 words (all levels), decoder blob.
 * Sizes in bytes via the size query with NULL buffers. */
int ntc_ctx_get_compressed(ntc_context* ctx, uint8_t* g0, size_t* g0_size,
                           uint8_t* g1, size_t* g1_size, uint8_t* decoder,
                           size_t* decoder_size, float* caldata4);
int ntc_ctx_dump(ntc_context* ctx, const char* path);
/* Applies the latent-grid Adam steps the fused path has deferred (the
 * zero-gradient steps of cells outside the recent tile footprints) to every
 * cell and waits for it. Reads and writes of the grids, Eval and the lock do
 * this by themselves; a caller timing training includes it so that all of the
 * work done is counted. No counterpart in the reference (it updates every
 * cell every step). */
int ntc_ctx_flush_grids(ntc_context* ctx);
/* Device pointer to the context's stream (for timing with events). */
void* ntc_ctx_stream(ntc_context* ctx);
/* Average durations over the last Train call with timing on, measured with
 * events: {fused kernel (the three launches of ntc_split_train_step where it
 * is used), decoder update (with its table refresh), grid update, steps
 * timed, grid updates timed, grid updates that skipped the zero gradient
 * outside the tile footprints, deferred catch-up, hot part (ms), steps with
 * the deferred update, MLP launch of ntc_split_train_step}; -1 when not
 * recorded. */
int ntc_ctx_kernel_times(ntc_context* ctx, float* out, int32_t n);
/* enabled: 0 off; 1 events around every kernel of each fused step; 2 around
 * the fused kernel only, or only around the MLP launch of the three-launch
 * step (fewer events between the step's kernels; the other figures of
 * ntc_ctx_kernel_times are then -1). Other values: InvalidState. */
int ntc_ctx_set_timing(ntc_context* ctx, int32_t enabled);
/* With timing 2: events only in the first of every `stride` steps of a Train
 * call (stride >= 1; 1, the default: every step), each event pair costing
 * the stream a few microseconds between two kernels; the averages of
 * ntc_ctx_kernel_times and its count of steps timed are over those steps. */
int ntc_ctx_set_timing_stride(ntc_context* ctx, int32_t stride);
/* Form of the fused path's step, taken at the next BeginTraining: 0 (the
 * default) the three launches of ntc_split_train_step where the layout has
 * them, else the one-launch ntc_fused_train_step; 1 always the one-launch
 * kernel. Other values: InvalidState. ntc_ctx_fused_form returns the form the
 * current training uses (0 three launches, 1 one launch). */
int ntc_ctx_set_fused_form(ntc_context* ctx, int32_t form);
int ntc_ctx_fused_form(ntc_context* ctx);
/* Fault injection for tests of the error paths. No counterpart in the
 * reference; it makes the paths where the reference returns CudaFailure or
 * AllocationFailure after a failed CUDA call (libopenntc.cpp:17-30, 506-527)
 * reachable on demand. Arms `site` so that its nth check from now on (counted
 * over every context of the process, n >= 1) fails; n = 0 disarms. Sites:
 * NTC_FAULT_ALLOC, a device allocation of a Context or Tensor (hipMalloc is
 * not called; out of memory); NTC_FAULT_LAUNCH, a kernel launch of a Context's
 * step, update, evaluation or export (the launch is skipped); NTC_FAULT_COPY,
 * a Context's blocking copy between host and device (skipped). Returns 0, or
 * NTC_E_INVALID for an unknown site or n < 0. */
#define NTC_FAULT_ALLOC 1
#define NTC_FAULT_LAUNCH 2
#define NTC_FAULT_COPY 3
int ntc_debug_inject_fault(int32_t site, int32_t nth);

/* ---- several texture sets trained together --------------------------------
 * No counterpart in the reference (one Context trains one set). The contexts
 * (each after BeginTrainingEx on the fused path, none trained yet, the same
 * decoder shape and step count, channel counts multiples of 4, at most 16
 * tiles per step) keep their own state; ntc_batch_train advances all of them
 * together, one launch per stage and step, optionally replaying one captured
 * hipGraph per step, and completes each context (lock, export, calibration)
 * like Context::Train. While the batch exists the contexts are trained only
 * through it. pass_texels / fused_grid 0: the fused launch's pass size and
 * workgroup budget are tuned by timing the first steps (cached per decoder
 * shape and set sizes for the process). */
typedef struct ntc_batch ntc_batch;

typedef struct ntc_batch_stats
{
  int32_t n_sets, steps, steps_done;
  int32_t pass_texels, fused_grid, grid_update_grid;
  int32_t tuned;        /* 1 tuned at creation, 2 from the cache, 0 given */
  int32_t graph;        /* 1 once a step graph is instantiated */
  int64_t texels_done;  /* texels of all sets over the steps run */
  float tune_ms[4];     /* fused launch time per step of each candidate */
  int32_t tune_pass[4], tune_grid[4];
} ntc_batch_stats;

int ntc_batch_create(ntc_context* const* sets, int32_t n_sets,
                     int32_t pass_texels, int32_t fused_grid, ntc_batch** out);
int ntc_batch_train(ntc_batch* b, int32_t n_steps, int32_t use_graph);
int ntc_batch_get_stats(const ntc_batch* b, ntc_batch_stats* out);
void ntc_batch_destroy(ntc_batch* b);

/* ---- texture manifest front end (Context::LoadManifest, libopenntc.cpp:
 * 1103-1207, 1296-1313), GPU-free parts -------------------------------------
 * ntc_manifest_parse checks a manifest with the reference's rules and returns
 * its sources in training order (stable sort by semantic), then decodes every
 * image; the status is what LoadManifest returns before any GPU work.
 * ntc_image_decode_rgba8 decodes one image (JPEG, PNG, BMP, TGA, binary PNM) to RGBA8
 * (rgba may be NULL to query the size). */
typedef struct ntc_manifest_source
{
  char name[64];
  char path[256];
  char channels[8];
  int32_t semantic;
  int32_t is_srgb, vertical_flip;
} ntc_manifest_source;

int ntc_manifest_parse(const char* path, int32_t min_dimension, int32_t* dim,
                       int32_t* count, ntc_manifest_source* sources /* 16 */);
int ntc_image_decode_rgba8(const char* path, int32_t* width, int32_t* height,
                           uint8_t* rgba, size_t capacity);

/* ---- synthetic PBR set ----------------------------------------------------
 * Seeded 9-channel set: albedo RGB (value noise through sRGB decode),
 * normal RGB (3-octave value-noise height field), roughness / metallic / AO
 * (independent value noise). Writes 3 RGBA8 images of dim*dim*4 bytes (host):
 * albedo, normal, orm (R = AO, G = roughness, B = metallic). */
int ntc_synthetic_pbr(int32_t dim, uint32_t seed, uint8_t* albedo,
                      uint8_t* normal, uint8_t* orm);

/* ---- ONTC container, GPU-free ---------------------------------------------
 * The reference's free functions Dump, Load, FillNTCConstants and
 * FileData::Data (openntc_runtime.h:123-150, openntc_runtime.cpp:156-468)
 * for FFI callers. ntc_compressed carries the fields of CompressedData
 * (openntc_runtime.h:91-121) as plain C types; its blob pointers are not
 * owned. Status values are openntc::Result. */
typedef struct ntc_compressed
{
  const void* g0[6];
  const void* g1[6];
  const void* decoder;
  uint64_t g0_size[6], g1_size[6], g0_offset[6], g1_offset[6];
  uint64_t g0_total_size, g1_total_size, decoder_size;
  int32_t g0_grid_dim[6], g1_grid_dim[6];
  int32_t g0_bits, g1_bits, g0_channels, g1_channels;
  int32_t dim, mip_count, level_count, channel_count;
  int32_t channel_semantics[16];
  int32_t profile;
  float max_abs_a1, max_abs_a2, s_a1, s_a2; /* CalibrationData */
} ntc_compressed;

typedef struct ntc_file ntc_file; /* openntc::FileData */

/* Dump (openntc_runtime.cpp:179-253). */
int ntc_container_dump(const char* path, const ntc_compressed* data);
/* Load (openntc_runtime.cpp:307-456): on success *out owns the file's bytes
 * and stays valid until ntc_file_free; on failure *out is NULL. */
int ntc_container_load(const char* path, ntc_file** out);
/* FileData::Data (openntc_runtime.cpp:465-468): pointers into the file. */
int ntc_file_data(const ntc_file* f, ntc_compressed* out);
void ntc_file_free(ntc_file* f);
/* FillNTCConstants (openntc_runtime.cpp:156-177) into 240 bytes. */
int ntc_fill_constants(const ntc_compressed* data, void* constants240);

#ifdef __cplusplus
}
#endif

#endif /* NTC_HIP_H */
