// openntc::Context implementation state (host side, C++). Shared by the
// Context methods (context.cpp) and the flat C shim (capi.cpp).
#pragma once

#include <libopenntc/libopenntc.h>

#include <hip/hip_runtime.h>

#include <random>
#include <string>
#include <vector>

#include "ntc_hip.h"

namespace openntc
{
// Distributions over std::mt19937 used for the parameter init and the step
// schedule. std::mt19937 itself is fully specified by the C++ standard; the
// reference's <random> distributions are not (libopenntc.cpp:137, 564-583),
// so these are defined here and restated by the CPU oracle.
float DrawUniform01(std::mt19937& g);       // [0,1), 24 random bits
float DrawUniform01Open0(std::mt19937& g);  // (0,1], 24 random bits
float DrawUniformRange(std::mt19937& g, float lo, float hi);
int32_t DrawUniformInt(std::mt19937& g, int32_t lo, int32_t hi);
bool DrawBernoulli(std::mt19937& g, double p);

struct StepPlan
{
  int32_t lod, level, tile_dim, n_tiles;
  int32_t origins[16];
  float lr_grid, lr_decoder;
  float bias1, bias2, grid_bias1, grid_bias2;
  int32_t grid_update, lock;
};

// Fault injection for tests of the error paths: the armed site's n-th check
// from now on fails as HIP would have (allocation: out of memory; launch: the
// launch is skipped and reported failed; copy: the copy is skipped and
// reported failed). Process-wide, one-shot per arming.
enum FaultSite : int
{
  kFaultAlloc = 1,
  kFaultLaunch = 2,
  kFaultCopy = 3,
};
bool FaultHit(int site);
void ArmFault(int site, int nth);

struct DeviceBuffer
{
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t Alloc(size_t n);
  void Free();
  template <typename T>
  T* as() const
  {
    return static_cast<T*>(ptr);
  }
};

struct Context::Impl
{
  // ---- configuration (Init / InitEx) ----
  bool initialized = false;
  bool manifest_loaded = false;
  Result last_error = Result::Success;
  TrainPhase phase = TrainPhase::TrainError;
  Profile profile = Profile::Bpp_0_2;
  int32_t g0_bits = 0, g1_bits = 0, g0_channels = 0, g1_channels = 0, g0_scale = 0;
  int32_t hidden_dim = 64;
  FeatureLayout feature_layout = FeatureLayout::Reference;
  Activation activation = Activation::HardGELU;
  int32_t feature_dim = 0, feature_dim_padded = 0, out_dim = 0;
  static constexpr int32_t kOutPadded = 16;
  // timing events per fused step: fused kernel start/end, after the decoder
  // update + table pack, grid catch-up start/end, end of the grid update,
  // MLP launch start/end of the three-launch step, decoder update start (on
  // the second stream with the three launches)
  static constexpr int32_t kStepEvents = 9;
  int32_t min_dimension = Context::kMinDimension;
  int32_t mip_min_dim = 4;
  uint32_t seed = 123;
  float g0_delta = 0.f, g1_delta = 0.f;
  std::mt19937 gen{123};

  // ---- manifest-dependent layout ----
  int32_t mip_count = 0, level_count = 0;
  int32_t mip_dim[kMaxMips] = {};
  int32_t g0_dim[kMaxLevels] = {}, g1_dim[kMaxLevels] = {};
  Manifest manifest = {};

  // ---- device state ----
  hipStream_t stream = nullptr;
  DeviceBuffer g0[kMaxLevels], g1[kMaxLevels];
  DeviceBuffer g0m[kMaxLevels], g0v[kMaxLevels], g1m[kMaxLevels], g1v[kMaxLevels];
  DeviceBuffer dg0[kMaxLevels], dg1[kMaxLevels];
  DeviceBuffer dec, dec_m, dec_v, dec_grad;  // W0[H][F] | W1[H][H] | Wout[16][H]
  DeviceBuffer package[kMaxMips];
  DeviceBuffer mips[kMaxSources][kMaxMips];
  DeviceBuffer origins_table;  // 16 ints per scheduled step
  DeviceBuffer eval_origins;   // 16 ints
  DeviceBuffer scalars;        // doubles: [0..15] channel SSE, [16] step SSE
  DeviceBuffer pack_scratch;
  // plain path activations (allocated on first use)
  int64_t plain_batch = 0;
  DeviceBuffer x, dx, yt, z0, a0, z1, a1, yp, dy, da1, dz1, da0, dz0, gemm_ws, loss_ws, chan_ws;
  DeviceBuffer decode_table, decode_ws;  // reduced-precision evaluation
  // Any change of latents or weights bumps state_version; decode_table holds
  // the table of decode_table_prec built at decode_table_version.
  uint64_t state_version = 0, decode_table_version = ~0ull;
  int32_t decode_table_prec = -1;

  // ---- training state ----
  int32_t batch_count = 0, batch_i = 0, lock_i = 0, grids_per_batch = 1;
  int32_t grid_batch_i[kMaxLevels] = {};
  float lr_grid_max = 0.01f, lr_dec_max = 0.005f;
  float init_lo = -0.1f, init_hi = 0.1f;
  TrainPath path = TrainPath::Fused;
  std::vector<StepPlan> plan;
  double last_step_sse = 0.0;

  // fused path state
  DeviceBuffer wtab;          // BF16 MFMA weight tables
  DeviceBuffer dw_part;       // [blocks][decoder elems]
  DeviceBuffer sse_part;      // [blocks]
  int32_t max_blocks = 0;
  // The step as three launches (gather, MLP, latent scatter; feature and dX
  // rows in split_ws) where compiled for the layout (ntc_split_supported),
  // unless fused_form = 1 asks for the one-launch kernel; decided at
  // BeginTraining.
  int32_t fused_form = 0;
  bool split = false;
  DeviceBuffer split_ws;
  // Deferred zero-gradient grid steps (fused path): per level, the count of
  // the level's Adam steps applied to every g0 / g1 cell, {lr, bias1, bias2}
  // of each of the level's steps, the steps issued so far and whether some
  // cell is behind them.
  bool lazy = false;
  DeviceBuffer lazy_last[kMaxLevels][2];
  DeviceBuffer lazy_steps[kMaxLevels];
  int32_t lazy_done[kMaxLevels] = {};
  bool lazy_behind[kMaxLevels] = {};
  // The next step's catch-up runs on side once ready_ev (after the step's
  // fused kernel) has fired, beside the step's decoder update, table pack and
  // hot part, whose cells it leaves out; caught_step: the step it was issued
  // for (-1: none), only ever the next step of the same Train call.
  // train_end: that call's end.
  hipStream_t side = nullptr;
  hipEvent_t ready_ev = nullptr, catch_ev = nullptr;
  int32_t caught_step = -1, train_end = 0;
  // work issued on side that the context stream has not waited for yet: the
  // next step's catch-up (catch_ev is recorded after it) and, with the three
  // launches, the step's decoder update (dec_ev; waited for before the next
  // MLP launch)
  bool side_pending = false, dec_pending = false;
  hipEvent_t dec_ev = nullptr;

  // ---- timing (optional) ----
  // When enabled, Train records events around the fused step kernel, the
  // decoder update (+ table pack) and the dense grid update of every step on
  // the context stream; kernel_ms holds {fused, decoder, grid, steps timed,
  // grid-update launches timed, grid updates that skipped the zero gradient
  // outside the tile footprints, catch-up, hot part, steps with the deferred
  // update} over the last Train call (durations averaged; with the deferred
  // update the grid figure is catch-up + hot), then the MLP launch of the
  // three-launch step (-1 for the one-launch kernel). Level 2 records only the
  // fused kernel's pair (each event recorded between the step's kernels delays
  // the next one by a few microseconds), or with the three launches only the
  // MLP launch's; the other figures are then -1.
  int32_t timing = 0;
  int32_t timing_stride = 1;  // timing 2: events in the first of every timing_stride steps
  int32_t tiles_update_steps = 0;
  std::vector<float> kernel_ms;
  std::vector<hipEvent_t> events;
  std::vector<uint8_t> ev_lazy;  // per timed step: deferred grid update (catch-up + hot)

  // ---- helpers ----
  int64_t DecoderElems() const
  {
    return (int64_t)hidden_dim * feature_dim + (int64_t)hidden_dim * hidden_dim + (int64_t)kOutPadded * hidden_dim;
  }
  int64_t GridElems(int which, int level) const
  {
    return which == 0 ? (int64_t)g0_dim[level] * g0_dim[level] * g0_channels
                      : (int64_t)g1_dim[level] * g1_dim[level] * g1_channels;
  }
  int32_t LevelForLod(int32_t lod) const;
  ntc_feature_desc FeatureDesc(int level) const;
  Result ApplyLayout(const LayoutInfo& l);
  Result ApplyProfile(Profile p);
  Result InitDevice();
  void FreeManifestBuffers();
  void FreeAll();
  Result SetupTextures(const TextureInput* tex, int32_t count, int32_t dim);
  Result EnsurePlainBuffers(int64_t batch);
  void BuildPlan(int32_t steps);
  Result RunPlainStep(const StepPlan& st, int32_t step_index);
  Result EnsureFusedBuffers();
  Result PackDecoderTables();
  Result RunFusedStep(const StepPlan& st, int32_t step_index, hipEvent_t* ev);
  int ActCode() const { return activation == Activation::HardGELU ? NTC_ACT_HARDGELU : NTC_ACT_RELU; }
  // Blocking copy ordered after all work queued on the context's stream. The
  // stream is non-blocking, so a plain hipMemcpy (legacy null stream) would
  // not wait for the kernels that produce or read the buffer.
  hipError_t CopySync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
  {
    if (FaultHit(kFaultCopy)) return hipErrorUnknown;
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
    return e == hipSuccess ? hipStreamSynchronize(stream) : e;
  }
  Result LockGrids();
  // Deferred grid steps: set-up at BeginTraining, and the catch-up of every
  // cell of every level before anything reads or writes the grids.
  Result SetupLazyGrids();
  Result FlushLazyGrids();
  Result Finalize(Context& ctx);
  Result ForwardTiles(int32_t mip, const int32_t* origins, int32_t n_tiles, int32_t T, bool with_targets);
  EvalResults Evaluate(const Context& ctx, EvalPrecision p, EvalDetail* detail, Result& err);
  // one eval batch, its tile origins already in device memory (dev_origins)
  Result DecodeTiles(int32_t mip, const int32_t* dev_origins, int32_t n_tiles, int32_t T, int32_t precision,
                     double* sse);
  CalibrationData RunCalibration(float headroom, Result& err);
};
} // namespace openntc

// The C ABI's context handle (include/ntc_hip.h).
struct ntc_context
{
  openntc::Context ctx;
};
