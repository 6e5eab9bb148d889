// Fused training step on BF16 MFMA (gfx950): the kernels and the C ABI of the
// training step (one set, several sets per launch), the decoder table packing
// and the decoder update. The step's body is in fused_body.h.
#include "decoder_update.h"

namespace
{
template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT, int NT>
__global__ __launch_bounds__(kThreads, 1) void k_fused_train(FusedParams P)
{
  fused_body<H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT, NT>(P, blockIdx.x, gridDim.x);
}

// Several independent training sets in one launch. For the step held in
// *step, set s owns workgroups [wg_first[s], wg_first[s + 1]) of the grid
// (wg_first has n_sets + 1 entries per step; workgroups past the last set's
// range have no work) and runs its batch with those as its grid.
template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT, int NT>
__global__ __launch_bounds__(kThreads, 1) void k_fused_train_sets(const FusedParams* __restrict__ params,
                                                                  const int32_t* __restrict__ wg_first,
                                                                  int n_sets, const int32_t* __restrict__ step)
{
  const int t = read_step(step);
  const int32_t* first = wg_first + (int64_t)t * (n_sets + 1);
  const int b = blockIdx.x;
  if (b >= first[n_sets]) return;
  int s = 0;
  while (s + 1 < n_sets && b >= first[s + 1]) ++s;
  const FusedParams& P = params[(int64_t)t * n_sets + s];
  fused_body<H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT, NT>(P, b - first[s], first[s + 1] - first[s]);
}

// ---- decoder table packing (FP32 masters -> permuted BF16 tables) ----------------
__host__ __device__ __forceinline__ void pack_table_elem(const TableLayout& L, int F, int O, int gslots,
                                                const float* __restrict__ dec, uint16_t* __restrict__ tab, int e)
{
  if (e >= L.total()) return;
  const int H = L.H;
  const float* w0 = dec;
  const float* w1 = dec + H * F;
  const float* wo = w1 + H * H;
  float v = 0.0f;
  if (e < L.oW1())
  {
    const int m = e / L.FP, c = swz(m, e % L.FP, L.FP);
    v = c < F ? w0[m * F + c] : 0.0f;
  }
  else if (e < L.oWo())
  {
    const int i = e - L.oW1(), m = i / H, pos = swz(m, i % H, H);
    v = w1[m * H + kperm_pos(pos)];
  }
  else
  {
    const int i = e - L.oWo(), o = i / H, pos = swz(o, i % H, H);
    v = o < O ? wo[o * H + kperm_pos(pos)] : 0.0f;
  }
  tab[e] = f2bf(v);
}

__global__ __launch_bounds__(256) void k_pack_tables(TableLayout L, int F, int O, int gslots,
                                                     const float* __restrict__ dec, uint16_t* __restrict__ tab)
{
  pack_table_elem(L, F, O, gslots, dec, tab, blockIdx.x * 256 + threadIdx.x);
}

// The tables of several sets of one shape: set blockIdx.y.
__global__ __launch_bounds__(256) void k_pack_tables_sets(TableLayout L, int F, int O, int gslots,
                                                          const ntc_pack_set* __restrict__ sets)
{
  const ntc_pack_set& S = sets[blockIdx.y];
  pack_table_elem(L, F, O, gslots, S.dec, S.table, blockIdx.x * 256 + threadIdx.x);
}

// ---- decoder update: ordered sum of workgroup partials + Adam ----------------------
__global__ __launch_bounds__(64) void k_decoder_update(int n_dec, int n_adam_w01, int H, int O, int blocks,
                                                        const float* __restrict__ part, DecAdam c,
                                                        float* __restrict__ grad, float* __restrict__ m,
                                                        float* __restrict__ v, float* __restrict__ p,
                                                        const double* __restrict__ sse_part,
                                                        double* __restrict__ sse_out, TableLayout L, int F,
                                                        int gslots, uint16_t* __restrict__ tab)
{
  decoder_update_elem(blockIdx.x * 64 + threadIdx.x, n_dec, n_adam_w01, H, O, blocks, part, c, grad, m, v, p,
                      sse_part, sse_out, L, F, gslots, tab);
}

// The decoder updates of several sets of one shape for the step held in
// *step: set blockIdx.y, its partial count and Adam scalars from
// steps[step][set].
__global__ __launch_bounds__(64) void k_decoder_update_sets(int n_dec, int n_adam_w01, int H, int O,
                                                             const ntc_dec_set* __restrict__ sets,
                                                             const ntc_dec_step* __restrict__ steps, int n_sets,
                                                             const int32_t* __restrict__ step)
{
  const int s = blockIdx.y;
  const ntc_dec_set& S = sets[s];
  const ntc_dec_step& st = steps[(int64_t)read_step(step) * n_sets + s];
  const DecAdam c = {st.lr, st.beta1, st.beta2, 1.0f - st.beta1, 1.0f - st.beta2, st.bias1, st.bias2};
  decoder_update_elem(blockIdx.x * 64 + threadIdx.x, n_dec, n_adam_w01, H, O, st.blocks, S.dw_partials, c,
                      S.grad, S.m, S.v, S.p, S.sse_partials, S.sse_out);
}

__global__ void k_step_advance(int32_t* step)
{
  if (threadIdx.x == 0) __hip_atomic_fetch_add(step, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- dispatch ---------------------------------------------------------------------------
typedef void (*FusedKernel)(FusedParams);
typedef void (*FusedSetsKernel)(const FusedParams*, const int32_t*, int, const int32_t*);

struct Entry
{
  Shape s;
  FusedKernel k;
  int pass;                // texels per workgroup pass
  FusedSetsKernel ks[2];   // multi-set variants: ks[0] with the pass above,
  int spass[2];            // ks[1] (if any) with a 64-texel pass
};

#define NTC_FUSED(H, FP, G, C0, C1, MODE, PEC, LODS, ACT)                                      \
  Entry{shape_v<H, FP, G, C0, C1, MODE, PEC, LODS, ACT>(),                                      \
        k_fused_train<H, FP, G, C0, C1, MODE, PEC, LODS, ACT, nt_for(H)>, 64 * nt_for(H),       \
        {k_fused_train_sets<H, FP, G, C0, C1, MODE, PEC, LODS, ACT, nt_for(H)>, nullptr},       \
        {64 * nt_for(H), 0}},
// ... and with a second tile shape (one 16-texel column tile per wave) for
// the multi-set launch's shape tuning
#define NTC_FUSED2(H, FP, G, C0, C1, MODE, PEC, LODS, ACT)                                     \
  Entry{shape_v<H, FP, G, C0, C1, MODE, PEC, LODS, ACT>(),                                      \
        k_fused_train<H, FP, G, C0, C1, MODE, PEC, LODS, ACT, nt_for(H)>, 64 * nt_for(H),       \
        {k_fused_train_sets<H, FP, G, C0, C1, MODE, PEC, LODS, ACT, nt_for(H)>,                 \
         k_fused_train_sets<H, FP, G, C0, C1, MODE, PEC, LODS, ACT, 1>},                        \
        {64 * nt_for(H), 64}},

const Entry kEntries[] = {NTC_FUSED_LAYOUTS(NTC_FUSED, NTC_FUSED2)};

const Entry* find_entry(const ntc_feature_desc& fd, int H, int act)
{
  Shape s;
  shape_of(fd, H, act, s);
  for (const Entry& e : kEntries)
    if (same(e.s, s)) return &e;
  return nullptr;
}
} // namespace

extern "C" {
int ntc_fused_supported(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t activation)
{
  if (!valid_feature_desc(fd)) return 0;
  return find_entry(*fd, hidden_dim, activation) != nullptr;
}

size_t ntc_fused_table_elems(const ntc_feature_desc* fd, int32_t hidden_dim)
{
  if (!valid_feature_desc(fd)) return 0;
  Shape s;
  shape_of(*fd, hidden_dim, 0, s);
  return (size_t)TableLayout{s.H, s.FP, s.G}.total();
}

int ntc_fused_max_blocks(int32_t* blocks)
{
  int dev = 0, cus = 0;
  NTC_HIP_TRY(hipGetDevice(&dev));
  NTC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (blocks) *blocks = cus;
  return NTC_OK;
}

int ntc_pack_decoder_bf16(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t out_dim, const float* dec,
                          uint16_t* table, hipStream_t stream)
{
  NTC_RETURN_IF(!valid_feature_desc(fd) || !dec || !table || out_dim < 1 || out_dim > 16, NTC_E_INVALID);
  NTC_RETURN_IF(hidden_dim % 64 != 0 || hidden_dim < 64 || hidden_dim > 256, NTC_E_INVALID);
  Shape s;
  shape_of(*fd, hidden_dim, 0, s);
  const TableLayout L{s.H, s.FP, s.G};
  const int gslots = g1_base(*fd) + fd->g1_channels;
  hipLaunchKernelGGL(k_pack_tables, dim3(blocks_for(L.total(), 256)), dim3(256), 0, stream, L, fd->feature_dim,
                     out_dim, gslots, dec, table);
  return NTC_LAUNCH_STATUS();
}

int ntc_fused_train_step(const ntc_fused_args* a, hipStream_t stream)
{
  NTC_RETURN_IF(!a || !valid_feature_desc(&a->feat) || !valid_tiles(&a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!pow2_tiles(a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!a->g0 || !a->g1 || !a->weights || !a->targets || !a->dw_partials || !a->sse_partials,
                NTC_E_INVALID);
  NTC_RETURN_IF(a->scatter && (!a->dg0 || !a->dg1), NTC_E_INVALID);
  NTC_RETURN_IF(a->out_dim < 1 || a->out_dim > 16 || a->blocks < 1, NTC_E_INVALID);
  const Entry* e = find_entry(a->feat, a->hidden_dim, a->activation);
  NTC_RETURN_IF(!e, NTC_E_INVALID);
  FusedParams P = {};
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
  P.B = batch_of(a->tiles);
  P.scatter = a->scatter;
  P.dg0 = a->dg0;
  P.dg1 = a->dg1;
  P.dw_part = a->dw_partials;
  P.sse_part = a->sse_partials;
  P.dec_elems = a->hidden_dim * a->feat.feature_dim + a->hidden_dim * a->hidden_dim + 16 * a->hidden_dim;
  const int64_t npass = (P.B + e->pass - 1) / e->pass;
  const int blocks = (int)(npass < a->blocks ? npass : a->blocks);
  NTC_RETURN_IF(blocks != a->blocks, NTC_E_INVALID);  // every partial slot must be written
  hipLaunchKernelGGL(e->k, dim3(blocks), dim3(kThreads), 0, stream, P);
  return NTC_LAUNCH_STATUS();
}

size_t ntc_fused_params_bytes(void) { return sizeof(FusedParams); }

int ntc_fused_params_pack(const ntc_fused_args* a, void* record)
{
  NTC_RETURN_IF(!a || !record || !valid_feature_desc(&a->feat) || !valid_tiles(&a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!pow2_tiles(a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!a->g0 || !a->g1 || !a->weights || !a->targets || !a->dw_partials || !a->sse_partials,
                NTC_E_INVALID);
  NTC_RETURN_IF(a->scatter && (!a->dg0 || !a->dg1), NTC_E_INVALID);
  NTC_RETURN_IF(a->out_dim < 1 || a->out_dim > 16, NTC_E_INVALID);
  NTC_RETURN_IF(!find_entry(a->feat, a->hidden_dim, a->activation), NTC_E_INVALID);
  FusedParams P = {};
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
  P.B = batch_of(a->tiles);
  P.scatter = a->scatter;
  P.dg0 = a->dg0;
  P.dg1 = a->dg1;
  P.dw_part = a->dw_partials;
  P.sse_part = a->sse_partials;
  P.dec_elems = a->hidden_dim * a->feat.feature_dim + a->hidden_dim * a->hidden_dim + 16 * a->hidden_dim;
  std::memcpy(record, &P, sizeof(P));
  return NTC_OK;
}

int ntc_fused_pass_sizes(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t activation, int32_t* sizes)
{
  if (!valid_feature_desc(fd) || !sizes) return 0;
  const Entry* e = find_entry(*fd, hidden_dim, activation);
  if (!e) return 0;
  int n = 0;
  for (int i = 0; i < 2; ++i)
    if (e->ks[i]) sizes[n++] = e->spass[i];
  return n;
}

int ntc_fused_train_sets(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t activation, int32_t pass_texels,
                         const void* params, const int32_t* wg_first, int32_t n_sets, const int32_t* step,
                         int32_t grid, hipStream_t stream)
{
  NTC_RETURN_IF(!valid_feature_desc(fd) || !params || !wg_first || !step || n_sets < 1 || grid < 1, NTC_E_INVALID);
  const Entry* e = find_entry(*fd, hidden_dim, activation);
  NTC_RETURN_IF(!e, NTC_E_INVALID);
  FusedSetsKernel k = nullptr;
  for (int i = 0; i < 2; ++i)
    if (e->ks[i] && e->spass[i] == pass_texels) k = e->ks[i];
  NTC_RETURN_IF(!k, NTC_E_INVALID);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, stream, static_cast<const FusedParams*>(params), wg_first,
                     n_sets, step);
  return NTC_LAUNCH_STATUS();
}

int ntc_decoder_update_sets(int32_t hidden_dim, int32_t feature_dim, int32_t out_dim, const ntc_dec_set* sets,
                            const ntc_dec_step* steps, int32_t n_sets, const int32_t* step, hipStream_t stream)
{
  NTC_RETURN_IF(hidden_dim < 1 || feature_dim < 1 || out_dim < 1 || out_dim > 16 || n_sets < 1, NTC_E_INVALID);
  NTC_RETURN_IF(!sets || !steps || !step, NTC_E_INVALID);
  const int H = hidden_dim;
  const int n_dec = H * feature_dim + H * H + 16 * H;
  hipLaunchKernelGGL(k_decoder_update_sets, dim3(blocks_for(n_dec, 64), n_sets), dim3(64), 0, stream, n_dec,
                     H * feature_dim + H * H, H, out_dim, sets, steps, n_sets, step);
  return NTC_LAUNCH_STATUS();
}

int ntc_pack_decoder_bf16_sets(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t out_dim,
                               const ntc_pack_set* sets, int32_t n_sets, hipStream_t stream)
{
  NTC_RETURN_IF(!valid_feature_desc(fd) || !sets || n_sets < 1 || out_dim < 1 || out_dim > 16, NTC_E_INVALID);
  NTC_RETURN_IF(hidden_dim % 64 != 0 || hidden_dim < 64 || hidden_dim > 256, NTC_E_INVALID);
  Shape s;
  shape_of(*fd, hidden_dim, 0, s);
  const TableLayout L{s.H, s.FP, s.G};
  const int gslots = g1_base(*fd) + fd->g1_channels;
  hipLaunchKernelGGL(k_pack_tables_sets, dim3(blocks_for(L.total(), 256), n_sets), dim3(256), 0, stream, L,
                     fd->feature_dim, out_dim, gslots, sets);
  return NTC_LAUNCH_STATUS();
}

int ntc_step_advance(int32_t* step, hipStream_t stream)
{
  NTC_RETURN_IF(!step, NTC_E_INVALID);
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, stream, step);
  return NTC_LAUNCH_STATUS();
}

int ntc_fused_blocks_for(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t activation, int64_t batch,
                         int32_t max_blocks)
{
  if (!valid_feature_desc(fd)) return 0;
  const Entry* e = find_entry(*fd, hidden_dim, activation);
  if (!e) return 0;
  const int64_t npass = (batch + e->pass - 1) / e->pass;
  return (int)(npass < max_blocks ? npass : max_blocks);
}

int ntc_decoder_update(int32_t hidden_dim, int32_t feature_dim, int32_t out_dim, int32_t blocks,
                       const float* dw_partials, float lr, float beta1, float beta2, float bias1, float bias2,
                       float* grad, float* m, float* v, float* p, const double* sse_partials, double* sse_out,
                       hipStream_t stream)
{
  NTC_RETURN_IF(hidden_dim < 1 || feature_dim < 1 || out_dim < 1 || out_dim > 16 || blocks < 1, NTC_E_INVALID);
  NTC_RETURN_IF(!dw_partials || !grad || !m || !v || !p, NTC_E_INVALID);
  const int H = hidden_dim;
  const int n_dec = H * feature_dim + H * H + 16 * H;
  DecAdam c = {lr, beta1, beta2, 1.0f - beta1, 1.0f - beta2, bias1, bias2};
  hipLaunchKernelGGL(k_decoder_update, dim3(blocks_for(n_dec, 64)), dim3(64), 0, stream, n_dec,
                     H * feature_dim + H * H, H, out_dim, blocks, dw_partials, c, grad, m, v, p, sse_partials,
                     sse_out, TableLayout{}, 0, 0, (uint16_t*)nullptr);
  return NTC_LAUNCH_STATUS();
}

int ntc_decoder_update_bf16(const ntc_feature_desc* fd, int32_t hidden_dim, int32_t out_dim, int32_t blocks,
                            const float* dw_partials, float lr, float beta1, float beta2, float bias1, float bias2,
                            float* grad, float* m, float* v, float* p, uint16_t* table,
                            const double* sse_partials, double* sse_out, hipStream_t stream)
{
  NTC_RETURN_IF(!fd || !valid_feature_desc(fd), NTC_E_INVALID);
  NTC_RETURN_IF(hidden_dim % 64 != 0 || hidden_dim < 64 || hidden_dim > 256, NTC_E_INVALID);
  NTC_RETURN_IF(out_dim < 1 || out_dim > 16 || blocks < 1, NTC_E_INVALID);
  NTC_RETURN_IF(!dw_partials || !grad || !m || !v || !p || !table, NTC_E_INVALID);
  Shape s;
  shape_of(*fd, hidden_dim, 0, s);
  const TableLayout L{s.H, s.FP, s.G};
  const int F = fd->feature_dim, H = hidden_dim;
  const int gslots = g1_base(*fd) + fd->g1_channels;
  const int n_dec = H * F + H * H + 16 * H;
  DecAdam c = {lr, beta1, beta2, 1.0f - beta1, 1.0f - beta2, bias1, bias2};
  hipLaunchKernelGGL(k_decoder_update, dim3(blocks_for(n_dec, 64)), dim3(64), 0, stream, n_dec, H * F + H * H, H,
                     out_dim, blocks, dw_partials, c, grad, m, v, p, sse_partials, sse_out, L, F, gslots, table);
  return NTC_LAUNCH_STATUS();
}
} // extern "C"
