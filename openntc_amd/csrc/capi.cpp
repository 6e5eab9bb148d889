// Flat C entry points over openntc::Context (declared in include/ntc_hip.h)
// for FFI callers, plus library-level queries. Status values are
// openntc::Result cast to int.
#include <cstring>
#include <string>

#include "context_impl.h"

using namespace openntc;

extern "C" {
const char* ntc_version(void) { return "openntc-mi355x 0.1 (gfx950)"; }

int ntc_dev_alloc(size_t bytes, void** out)
{
  if (!out) return NTC_E_INVALID;
  *out = nullptr;
  const hipError_t e = hipMalloc(out, bytes ? bytes : 16);
  if (e != hipSuccess) return (int)e;
  return (int)hipMemset(*out, 0, bytes ? bytes : 16);
}

int ntc_dev_free(void* p) { return p ? (int)hipFree(p) : NTC_OK; }

int ntc_dev_upload(void* dst, const void* src, size_t bytes)
{
  if (!bytes) return NTC_OK;
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return (int)e;
  return (int)hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

int ntc_dev_download(void* dst, const void* src, size_t bytes)
{
  if (!bytes) return NTC_OK;
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return (int)e;
  return (int)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
}

int ntc_dev_synchronize(void) { return (int)hipDeviceSynchronize(); }

int ntc_dev_copy(void* dst, const void* src, size_t bytes)
{
  if (!bytes) return NTC_OK;
  return (int)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
}

int ntc_device_check(int32_t* count)
{
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = e == hipSuccess ? n : 0;
  if (e != hipSuccess || n < 1) return NTC_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return NTC_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return NTC_E_NODEVICE;
  return NTC_OK;
}

const char* ntc_status_string(int status)
{
  switch (status)
  {
  case NTC_OK: return "ok";
  case NTC_E_INVALID: return "invalid argument";
  case NTC_E_WORKSPACE: return "workspace too small";
  case NTC_E_NODEVICE: return "no gfx950 device";
  default: return status > 0 ? hipGetErrorString((hipError_t)status) : "unknown";
  }
}

ntc_context* ntc_ctx_create(void) { return new ntc_context(); }

void ntc_ctx_destroy(ntc_context* c) { delete c; }

int ntc_ctx_init(ntc_context* c, int32_t profile)
{
  if (!c) return (int)Result::InvalidState;
  ContextInitInfo info = {};
  info.profile_ = (Profile)profile;
  return (int)c->ctx.Init(info);
}

int ntc_ctx_init_ex(ntc_context* c, const ntc_init_ex* in)
{
  if (!c || !in) return (int)Result::InvalidState;
  ContextInitInfoEx info = {};
  info.profile_ = (Profile)in->profile;
  info.use_layout_ = in->use_layout != 0;
  info.layout_.g0_bits_ = in->layout.g0_bits;
  info.layout_.g1_bits_ = in->layout.g1_bits;
  info.layout_.g0_channels_ = in->layout.g0_channels;
  info.layout_.g1_channels_ = in->layout.g1_channels;
  info.layout_.g0_scale_ = in->layout.g0_scale;
  info.layout_.hidden_dim_ = in->layout.hidden_dim;
  info.layout_.feature_layout_ = (FeatureLayout)in->layout.feature_layout;
  info.layout_.activation_ = (Activation)in->layout.activation;
  info.min_dimension_ = in->min_dimension;
  info.mip_min_dim_ = in->mip_min_dim;
  info.seed_ = in->seed;
  return (int)c->ctx.InitEx(info);
}

int ntc_ctx_load_textures(ntc_context* c, int32_t count, const uint8_t* const* rgba8, const int32_t* semantics,
                          const char* const* channels, int32_t dim)
{
  if (!c || !rgba8 || !semantics || !channels || count < 1 || count > kMaxSources) return (int)Result::InvalidManifest;
  TextureInput tex[kMaxSources] = {};
  for (int i = 0; i < count; ++i)
  {
    tex[i].rgba8_ = rgba8[i];
    tex[i].semantic_ = (Semantic)semantics[i];
    tex[i].channels_ = channels[i];
    tex[i].name_ = "";
  }
  return (int)c->ctx.LoadTextures(tex, count, dim);
}

int ntc_ctx_load_manifest(ntc_context* c, const char* path)
{
  if (!c || !path) return (int)Result::FileNotFound;
  return (int)c->ctx.LoadManifest(path);
}

void ntc_ctx_unload_manifest(ntc_context* c)
{
  if (c) c->ctx.UnloadManifest();
}

int ntc_ctx_begin_training(ntc_context* c, int32_t quality, int32_t grids_per_batch)
{
  if (!c) return (int)Result::InvalidState;
  TrainInfo info = {};
  info.quality_ = (Quality)quality;
  info.grids_per_batch_ = grids_per_batch;
  c->ctx.GetImpl()->last_error = Result::Success;
  c->ctx.BeginTraining(info);
  return (int)c->ctx.GetLastError();
}

int ntc_ctx_begin_training_ex(ntc_context* c, const ntc_train_ex* in)
{
  if (!c || !in) return (int)Result::InvalidState;
  TrainInfoEx info = {};
  info.steps_ = in->steps;
  info.grids_per_batch_ = in->grids_per_batch;
  info.lr_grid_ = in->lr_grid;
  info.lr_decoder_ = in->lr_decoder;
  info.init_lo_ = in->init_lo;
  info.init_hi_ = in->init_hi;
  info.path_ = (TrainPath)in->path;
  c->ctx.GetImpl()->last_error = Result::Success;
  c->ctx.BeginTrainingEx(info);
  return (int)c->ctx.GetLastError();
}

static void fill_progress(const TrainProgress& p, ntc_progress* out)
{
  if (!out) return;
  out->phase = (int32_t)p.phase_;
  out->result = (int32_t)p.result_;
  out->batches_complete = p.batches_complete_;
  out->total_batches = p.total_batches_;
}

int ntc_ctx_train(ntc_context* c, int32_t n, ntc_progress* out)
{
  if (!c) return (int)Result::InvalidState;
  const TrainProgress p = c->ctx.Train(n);
  fill_progress(p, out);
  return (int)p.result_;
}

int ntc_ctx_train_until_complete(ntc_context* c, ntc_progress* out)
{
  if (!c) return (int)Result::InvalidState;
  const TrainProgress p = c->ctx.TrainUntilComplete();
  fill_progress(p, out);
  return (int)p.result_;
}

int ntc_ctx_eval(ntc_context* c, int32_t precision, double* mse, double* psnr, double* channel_sse)
{
  if (!c) return (int)Result::InvalidState;
  c->ctx.GetImpl()->last_error = Result::Success;
  EvalDetail detail = {};
  const EvalResults r = c->ctx.EvalEx((EvalPrecision)precision, &detail);
  if (mse) *mse = r.mse_;
  if (psnr) *psnr = r.psnr_;
  if (channel_sse)
    for (int i = 0; i < kMaxChannels; ++i) channel_sse[i] = detail.sse_[i];
  return (int)c->ctx.GetLastError();
}

int ntc_ctx_calibrate(ntc_context* c, float headroom, float* out4)
{
  if (!c) return (int)Result::InvalidState;
  c->ctx.GetImpl()->last_error = Result::Success;
  const CalibrationData d = c->ctx.Calibrate(headroom);
  if (out4)
  {
    out4[0] = d.max_abs_a1_;
    out4[1] = d.max_abs_a2_;
    out4[2] = d.s_a1_;
    out4[3] = d.s_a2_;
  }
  return (int)c->ctx.GetLastError();
}

int ntc_ctx_get_last_error(ntc_context* c) { return c ? (int)c->ctx.GetLastError() : (int)Result::InvalidState; }

int ntc_ctx_get_layout(ntc_context* c, ntc_ctx_layout* out)
{
  if (!c || !out) return (int)Result::InvalidState;
  const Context::Impl& I = *c->ctx.GetImpl();
  std::memset(out, 0, sizeof(*out));
  out->dim = I.mip_dim[0];
  out->mip_count = I.mip_count;
  out->level_count = I.level_count;
  out->out_dim = I.out_dim;
  for (int i = 0; i < 12; ++i) out->mip_dim[i] = I.mip_dim[i];
  for (int i = 0; i < 6; ++i)
  {
    out->g0_dim[i] = I.g0_dim[i];
    out->g1_dim[i] = I.g1_dim[i];
  }
  out->g0_channels = I.g0_channels;
  out->g1_channels = I.g1_channels;
  out->g0_bits = I.g0_bits;
  out->g1_bits = I.g1_bits;
  out->g0_scale = I.g0_scale;
  out->feature_dim = I.feature_dim;
  out->feature_dim_padded = I.feature_dim_padded;
  out->hidden_dim = I.hidden_dim;
  out->out_dim_padded = Context::Impl::kOutPadded;
  out->activation = I.activation == Activation::HardGELU ? NTC_ACT_HARDGELU : NTC_ACT_RELU;
  if (I.level_count > 0) out->feat = I.FeatureDesc(0);
  out->g0_delta = I.g0_delta;
  out->g1_delta = I.g1_delta;
  out->lock_step = I.lock_i;
  out->total_steps = I.batch_count;
  out->grids_per_batch = I.grids_per_batch;
  out->batch_index = I.batch_i;
  out->seed = I.seed;
  return (int)(I.initialized ? Result::Success : Result::InvalidState);
}

int ntc_ctx_step_info(ntc_context* c, int32_t step, ntc_step_info* out)
{
  if (!c || !out) return (int)Result::InvalidState;
  const Context::Impl& I = *c->ctx.GetImpl();
  if (step < 0 || step >= (int)I.plan.size()) return (int)Result::InvalidState;
  const StepPlan& s = I.plan[step];
  out->lod = s.lod;
  out->level = s.level;
  out->tile_dim = s.tile_dim;
  out->n_tiles = s.n_tiles;
  std::memcpy(out->origins, s.origins, sizeof(s.origins));
  out->lr_grid = s.lr_grid;
  out->lr_decoder = s.lr_decoder;
  out->bias1 = s.bias1;
  out->bias2 = s.bias2;
  out->grid_bias1 = s.grid_bias1;
  out->grid_bias2 = s.grid_bias2;
  out->grid_update = s.grid_update;
  out->lock = s.lock;
  return (int)Result::Success;
}

int ntc_ctx_feature_level_for_lod(ntc_context* c, int32_t lod) { return c ? c->ctx.FeatureLevelForLod(lod) : -1; }

static DeviceBuffer* grid_buffer(Context::Impl& I, int which, int level)
{
  if (level < 0 || level >= I.level_count) return nullptr;
  switch (which)
  {
  case 0: return &I.g0[level];
  case 1: return &I.g1[level];
  case 2: return &I.g0m[level];
  case 3: return &I.g0v[level];
  case 4: return &I.g1m[level];
  case 5: return &I.g1v[level];
  case 6: return &I.dg0[level];
  case 7: return &I.dg1[level];
  default: return nullptr;
  }
}

int ntc_ctx_read_grid(ntc_context* c, int32_t which, int32_t level, float* host)
{
  if (!c || !host) return (int)Result::InvalidState;
  Context::Impl& I = *c->ctx.GetImpl();
  if (!I.manifest_loaded) return (int)Result::InvalidState;
  DeviceBuffer* b = grid_buffer(I, which, level);
  if (!b || !b->ptr) return (int)Result::InvalidState;
  // every cell current: the deferred zero-gradient steps applied
  if (I.FlushLazyGrids() != Result::Success) return (int)Result::CudaFailure;
  if (hipStreamSynchronize(I.stream) != hipSuccess) return (int)Result::CudaFailure;
  return hipMemcpy(host, b->ptr, b->bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : (int)Result::CudaFailure;
}

int ntc_ctx_write_grid(ntc_context* c, int32_t which, int32_t level, const float* host)
{
  if (!c || !host) return (int)Result::InvalidState;
  Context::Impl& I = *c->ctx.GetImpl();
  if (!I.manifest_loaded) return (int)Result::InvalidState;
  DeviceBuffer* b = grid_buffer(I, which, level);
  if (!b || !b->ptr) return (int)Result::InvalidState;
  // every cell current: the deferred zero-gradient steps applied
  if (I.FlushLazyGrids() != Result::Success) return (int)Result::CudaFailure;
  if (hipStreamSynchronize(I.stream) != hipSuccess) return (int)Result::CudaFailure;
  ++I.state_version;
  return hipMemcpy(b->ptr, host, b->bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : (int)Result::CudaFailure;
}

static DeviceBuffer* decoder_buffer(Context::Impl& I, int which)
{
  switch (which)
  {
  case 0: return &I.dec;
  case 1: return &I.dec_m;
  case 2: return &I.dec_v;
  case 3: return &I.dec_grad;
  default: return nullptr;
  }
}

int ntc_ctx_read_decoder(ntc_context* c, int32_t which, float* host)
{
  if (!c || !host) return (int)Result::InvalidState;
  Context::Impl& I = *c->ctx.GetImpl();
  DeviceBuffer* b = decoder_buffer(I, which);
  if (!I.initialized || !b || !b->ptr) return (int)Result::InvalidState;
  if (hipStreamSynchronize(I.stream) != hipSuccess) return (int)Result::CudaFailure;
  return hipMemcpy(host, b->ptr, b->bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : (int)Result::CudaFailure;
}

int ntc_ctx_write_decoder(ntc_context* c, int32_t which, const float* host)
{
  if (!c || !host) return (int)Result::InvalidState;
  Context::Impl& I = *c->ctx.GetImpl();
  DeviceBuffer* b = decoder_buffer(I, which);
  if (!I.initialized || !b || !b->ptr) return (int)Result::InvalidState;
  if (hipStreamSynchronize(I.stream) != hipSuccess) return (int)Result::CudaFailure;
  ++I.state_version;
  return hipMemcpy(b->ptr, host, b->bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : (int)Result::CudaFailure;
}

int ntc_ctx_read_package(ntc_context* c, int32_t mip, uint8_t* host)
{
  if (!c || !host) return (int)Result::InvalidState;
  Context::Impl& I = *c->ctx.GetImpl();
  if (!I.manifest_loaded || mip < 0 || mip >= I.mip_count) return (int)Result::InvalidState;
  if (hipStreamSynchronize(I.stream) != hipSuccess) return (int)Result::CudaFailure;
  return hipMemcpy(host, I.package[mip].ptr, I.package[mip].bytes, hipMemcpyDeviceToHost) == hipSuccess
             ? 0
             : (int)Result::CudaFailure;
}

int ntc_ctx_get_compressed(ntc_context* c, uint8_t* g0, size_t* g0_size, uint8_t* g1, size_t* g1_size,
                           uint8_t* decoder, size_t* decoder_size, float* caldata4)
{
  if (!c) return (int)Result::InvalidState;
  c->ctx.GetImpl()->last_error = Result::Success;
  const CompressedData d = c->ctx.GetCompressedData();
  if (c->ctx.GetLastError() != Result::Success) return (int)c->ctx.GetLastError();
  if (g0_size) *g0_size = d.g0_total_size_;
  if (g1_size) *g1_size = d.g1_total_size_;
  if (decoder_size) *decoder_size = d.decoder_size_;
  if (g0)
    for (int i = 0; i < d.level_count_; ++i) std::memcpy(g0 + d.g0_offset_[i], d.g0_[i], d.g0_size_[i]);
  if (g1)
    for (int i = 0; i < d.level_count_; ++i) std::memcpy(g1 + d.g1_offset_[i], d.g1_[i], d.g1_size_[i]);
  if (decoder) std::memcpy(decoder, d.decoder_, d.decoder_size_);
  if (caldata4)
  {
    caldata4[0] = d.caldata_.max_abs_a1_;
    caldata4[1] = d.caldata_.max_abs_a2_;
    caldata4[2] = d.caldata_.s_a1_;
    caldata4[3] = d.caldata_.s_a2_;
  }
  return 0;
}

int ntc_ctx_dump(ntc_context* c, const char* path)
{
  if (!c || !path) return (int)Result::InvalidState;
  c->ctx.GetImpl()->last_error = Result::Success;
  const CompressedData d = c->ctx.GetCompressedData();
  if (c->ctx.GetLastError() != Result::Success) return (int)c->ctx.GetLastError();
  return (int)Dump(path, d);
}

void* ntc_ctx_stream(ntc_context* c) { return c ? (void*)c->ctx.GetImpl()->stream : nullptr; }

int ntc_ctx_kernel_times(ntc_context* c, float* out, int32_t n)
{
  if (!c || !out) return (int)Result::InvalidState;
  const auto& v = c->ctx.GetImpl()->kernel_ms;
  for (int i = 0; i < n; ++i) out[i] = i < (int)v.size() ? v[i] : -1.0f;
  return 0;
}

int ntc_ctx_flush_grids(ntc_context* c)
{
  if (!c) return (int)Result::InvalidState;
  Context::Impl& I = *c->ctx.GetImpl();
  if (!I.manifest_loaded) return (int)Result::InvalidState;
  if (I.FlushLazyGrids() != Result::Success || hipStreamSynchronize(I.stream) != hipSuccess)
    return (int)Result::CudaFailure;
  return 0;
}

int ntc_ctx_set_timing(ntc_context* c, int32_t enabled)
{
  if (!c) return (int)Result::InvalidState;
  if (enabled < 0 || enabled > 2) return (int)Result::InvalidState;
  c->ctx.GetImpl()->timing = enabled;
  return 0;
}

int ntc_ctx_set_timing_stride(ntc_context* c, int32_t stride)
{
  if (!c) return (int)Result::InvalidState;
  if (stride < 1) return (int)Result::InvalidState;
  c->ctx.GetImpl()->timing_stride = stride;
  return 0;
}

int ntc_ctx_set_fused_form(ntc_context* c, int32_t form)
{
  if (!c) return (int)Result::InvalidState;
  if (form < 0 || form > 1) return (int)Result::InvalidState;
  c->ctx.GetImpl()->fused_form = form;
  return 0;
}

int ntc_ctx_fused_form(ntc_context* c)
{
  if (!c) return -1;
  const Context::Impl& I = *c->ctx.GetImpl();
  return I.split ? 0 : 1;
}

int ntc_debug_inject_fault(int32_t site, int32_t nth)
{
  if (site < NTC_FAULT_ALLOC || site > NTC_FAULT_COPY || nth < 0) return NTC_E_INVALID;
  openntc::ArmFault(site, nth);
  return 0;
}
} // extern "C"
