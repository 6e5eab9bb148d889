// The fused step as a forward / backward pair (BF16 MFMA, FP32 accumulation,
// FP32 master state). The reference splits its step at the loss: the forward
// is draw_features + three GEMMs + activations (libopenntc.cpp:621-655), the
// loss stage forms dL/dYhat (:659-668), the backward is six GEMMs +
// activation backward + accumulate_grid_gradients (:672-709,
// ntc_kernel.cu:261-373). Here both halves are launches of the fused body
// (fused_body.h): the forward writes Yhat, the backward takes an upstream
// dL/dYhat. Recompute policy: nothing is kept between the two calls; the
// backward gathers and runs the forward again, which costs a third of the
// step's MFMA work and no HBM for activations.
#include <mutex>

#include "fused_body.h"

namespace
{
template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT, int NT>
__global__ __launch_bounds__(kThreads, 1) void k_fused_forward(FusedParams P)
{
  fused_body<H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT, NT, kForward>(P, blockIdx.x, gridDim.x);
}

template <int H, int FP, int G, int C0, int C1, int G0MODE, int PEC, int LODS, int ACT, int NT>
__global__ __launch_bounds__(kThreads, 1) void k_fused_backward(FusedParams P)
{
  fused_body<H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT, NT, kBackward>(P, blockIdx.x, gridDim.x);
}

// Ordered sum of the workgroup partials (the same block order, and so the
// same bits, as the decoder update's gradient) without the Adam step.
__global__ __launch_bounds__(64) void k_reduce_partials(int n_dec, int blocks, const float* __restrict__ part,
                                                        float* __restrict__ grad, const double* __restrict__ sse_part,
                                                        double* __restrict__ sse_out)
{
  const int e = blockIdx.x * 64 + threadIdx.x;
  if (e == 0 && sse_out)
  {
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += sse_part[b];
    *sse_out = s;
  }
  if (e >= n_dec) return;
  float gsum = 0.0f;
  int b = 0;
  for (; b + 32 <= blocks; b += 32)
  {
    float v[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = part[(int64_t)(b + k) * n_dec + e];
#pragma unroll
    for (int k = 0; k < 32; ++k) gsum = gsum + v[k];
  }
  for (; b < blocks; ++b) gsum = gsum + part[(int64_t)b * n_dec + e];
  grad[e] = gsum;
}

typedef void (*PairKernel)(FusedParams);

struct PairEntry
{
  Shape s;
  PairKernel fwd, bwd;
  int pass;
};

#define NTC_PAIR(H, FP, G, C0, C1, MODE, PEC, LODS, ACT)                                   \
  PairEntry{shape_v<H, FP, G, C0, C1, MODE, PEC, LODS, ACT>(),                              \
            k_fused_forward<H, FP, G, C0, C1, MODE, PEC, LODS, ACT, nt_for(H)>,             \
            k_fused_backward<H, FP, G, C0, C1, MODE, PEC, LODS, ACT, nt_for(H)>, 64 * nt_for(H)},

const PairEntry kPairs[] = {NTC_FUSED_LAYOUTS(NTC_PAIR, NTC_PAIR)};

const PairEntry* find_pair(const ntc_feature_desc& fd, int H, int act)
{
  Shape s;
  shape_of(fd, H, act, s);
  for (const PairEntry& e : kPairs)
    if (same(e.s, s)) return &e;
  return nullptr;
}

// Argument checks shared by both halves; fills the body's parameters.
int pair_params(const ntc_fused_args* a, FusedParams& P, const PairEntry*& e)
{
  NTC_RETURN_IF(!a || !valid_feature_desc(&a->feat) || !valid_tiles(&a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!pow2_tiles(a->tiles), NTC_E_INVALID);
  NTC_RETURN_IF(!a->g0 || !a->g1 || !a->weights, NTC_E_INVALID);
  NTC_RETURN_IF(a->out_dim < 1 || a->out_dim > 16 || a->blocks < 1, NTC_E_INVALID);
  e = find_pair(a->feat, a->hidden_dim, a->activation);
  NTC_RETURN_IF(!e, NTC_E_INVALID);
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
  return NTC_OK;
}
} // namespace

extern "C" {
int ntc_step_forward(const ntc_fused_args* a, float* yhat, hipStream_t stream)
{
  FusedParams P;
  const PairEntry* e = nullptr;
  const int st = pair_params(a, P, e);
  if (st != NTC_OK) return st;
  NTC_RETURN_IF(!yhat, NTC_E_INVALID);
  NTC_RETURN_IF(a->targets && !a->sse_partials, NTC_E_INVALID);
  P.yhat = yhat;
  P.scatter = 0;
  hipLaunchKernelGGL(e->fwd, dim3(a->blocks), dim3(kThreads), 0, stream, P);
  return NTC_LAUNCH_STATUS();
}

int ntc_step_backward(const ntc_fused_args* a, const float* dyhat, hipStream_t stream)
{
  FusedParams P;
  const PairEntry* e = nullptr;
  const int st = pair_params(a, P, e);
  if (st != NTC_OK) return st;
  NTC_RETURN_IF(!dyhat || !a->dw_partials || !a->sse_partials, NTC_E_INVALID);
  NTC_RETURN_IF(a->scatter && (!a->dg0 || !a->dg1), NTC_E_INVALID);
  P.dyhat = dyhat;
  P.pkg = nullptr;
  hipLaunchKernelGGL(e->bwd, dim3(a->blocks), dim3(kThreads), 0, stream, P);
  return NTC_LAUNCH_STATUS();
}

int ntc_reduce_decoder_partials(int32_t hidden_dim, int32_t feature_dim, int32_t blocks, const float* dw_partials,
                                float* grad, const double* sse_partials, double* sse_out, hipStream_t stream)
{
  NTC_RETURN_IF(hidden_dim < 1 || feature_dim < 1 || blocks < 1 || !dw_partials || !grad, NTC_E_INVALID);
  NTC_RETURN_IF(sse_out && !sse_partials, NTC_E_INVALID);
  const int n_dec = hidden_dim * feature_dim + hidden_dim * hidden_dim + 16 * hidden_dim;
  hipLaunchKernelGGL(k_reduce_partials, dim3(blocks_for(n_dec, 64)), dim3(64), 0, stream, n_dec, blocks,
                     dw_partials, grad, sse_partials, sse_out);
  return NTC_LAUNCH_STATUS();
}
} // extern "C"

// ---- the training step without a workspace argument ----------------------------------
// For callers that cannot allocate device memory. The workgroup partials of
// the decoder gradient and of the SSE live in one file-scope device array,
// one copy per GPU, held until the process ends: kNoAllocFloats floats of
// partials (256 workgroups x 32768 decoder elements, 32 MiB), then
// kNoAllocBlocks doubles. Every call writes each partial slot it reads in
// full, so nothing is cleared between calls; calls that overlap on one GPU
// share the array and race.
namespace
{
constexpr int64_t kNoAllocFloats = (int64_t)256 * 32768;
constexpr int kNoAllocBlocks = 256;
static __device__ __align__(256) float scratch[kNoAllocFloats + 2 * kNoAllocBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current GPU's copy of the array, looked up once
// per device (the array's name is only a host-side stand-in).
int scratch_for_current_device(void** out)
{
  int dev = 0;
  NTC_HIP_TRY(hipGetDevice(&dev));
  NTC_RETURN_IF(dev < 0 || dev >= kMaxDevices, NTC_E_INVALID);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev])
  {
    void* ptr = nullptr;
    NTC_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    NTC_RETURN_IF(!ptr, NTC_E_INVALID);
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return NTC_OK;
}
} // namespace

extern "C" {
size_t ntc_fused_train_step_noalloc_max(void) { return (size_t)kNoAllocFloats; }

int ntc_fused_train_step_noalloc(const ntc_fused_args* a, float* grad, double* sse_out, hipStream_t stream)
{
  NTC_RETURN_IF(!a || !grad, NTC_E_INVALID);
  NTC_RETURN_IF(a->dw_partials || a->sse_partials, NTC_E_INVALID);  // the partials are internal here
  NTC_RETURN_IF(a->hidden_dim < 1 || a->feat.feature_dim < 1, NTC_E_INVALID);
  const int64_t dec = (int64_t)a->hidden_dim * a->feat.feature_dim + (int64_t)a->hidden_dim * a->hidden_dim +
                      16 * (int64_t)a->hidden_dim;
  NTC_RETURN_IF(a->blocks < 1 || a->blocks > kNoAllocBlocks, NTC_E_WORKSPACE);
  NTC_RETURN_IF((int64_t)a->blocks * dec > kNoAllocFloats, NTC_E_WORKSPACE);
  void* base = nullptr;
  const int st = scratch_for_current_device(&base);
  if (st != NTC_OK) return st;
  float* part = static_cast<float*>(base);
  double* ssep = reinterpret_cast<double*>(part + kNoAllocFloats);
  ntc_fused_args b = *a;
  b.dw_partials = part;
  b.sse_partials = ssep;
  const int s1 = ntc_fused_train_step(&b, stream);
  if (s1 != NTC_OK) return s1;
  return ntc_reduce_decoder_partials(a->hidden_dim, a->feat.feature_dim, a->blocks, part, grad, ssep, sse_out,
                                     stream);
}
} // extern "C"
