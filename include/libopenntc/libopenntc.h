// libopenntc public C++ API, MI355X implementation.
//
// Drop-in boundary: the Context class keeps the reference's method names,
// signatures, public fields, constants and state machine
// (libopenntc/libopenntc.h:8-332) so existing callers build unchanged. The
// reference's private members (CUDA tensors, cuBLAS handle, curand states) are
// replaced by one opaque implementation pointer; the class is therefore
// source-compatible, not binary-compatible, with the reference.
//
// Members marked "extension" have no counterpart in the reference. They exist
// for the measured configurations: arbitrary latent layouts, an explicit step
// count and learning rates, in-memory textures below 1024 px, and eval in
// several decode precisions.
#pragma once

#include <libopenntc/openntc_runtime.h>

#include <atomic>
#include <random>

namespace openntc
{
struct ContextInitInfo
{
  Profile profile_;
};

// Owning device buffers, row-major (libopenntc/libopenntc.h:13-91). Kept for
// API completeness; backed by hipMalloc.
class Tensor2d
{
public:
  Tensor2d();
  ~Tensor2d();

  Result Init(int x, int y);
  Result InitLike(const Tensor2d& t0);
  void FillZero();
  void FillKaiming(std::mt19937& gen);
  void Destroy();
  float* DevicePtr();
  size_t SizeBytes() const;
  size_t NumElems() const;

private:
  bool initialized_;
  int shape_[2];
  float* dev_;
};

class Tensor3d
{
public:
  Tensor3d();
  ~Tensor3d();

  Result Init(int x, int y, int z);
  Result InitLike(const Tensor3d& t0);
  void FillZero();
  void FillUniform(std::mt19937& gen, float lb, float ub);
  void Destroy();
  float* DevicePtr();
  bool IsInitialized() const;
  size_t SizeBytes() const;
  size_t NumElems() const;

private:
  bool initialized_;
  int shape_[3];
  float* dev_;
};

class U8Tensor3d
{
public:
  U8Tensor3d();
  ~U8Tensor3d();

  Result Init(int x, int y, int z);
  Result InitLike(const U8Tensor3d& t0);
  void FillZero();
  void Destroy();
  uint8_t* DevicePtr();
  bool IsInitialized() const;
  size_t SizeBytes() const;
  size_t NumElems() const;

private:
  bool initialized_;
  int shape_[3];
  uint8_t* dev_;
};

class IntTensor1d
{
public:
  IntTensor1d();
  ~IntTensor1d();

  Result Init(int x);
  void Destroy();
  int* DevicePtr();

private:
  bool initialized_;
  int shape_[1];
  int* dev_;
};

struct EvalResults
{
  double mse_;
  double psnr_;
};

enum TrainPhase : int32_t
{
  ManifestLoaded = 0,
  TrainInProgress = 1,
  TrainComplete = 2,

  TrainError = -1,
};

struct TrainProgress
{
  TrainPhase phase_;
  Result result_;
  int batches_complete_;
  int total_batches_;
};

enum class Channel : int32_t
{
  R = 0,
  G = 1,
  B = 2,
  A = 3,

  Count,
  Invalid = -1
};

struct TextureSource
{
  std::string path_;
  std::string name_;
  bool is_srgb_;
  Semantic semantic_;
  bool vertical_flip_;
  int32_t num_channels_;
  Channel channel_mapping_[4];
};

constexpr int32_t kMaxSources = 16;

struct Manifest
{
  TextureSource sources_[kMaxSources];
  int32_t source_count_;
  int32_t dim_;
};

enum class Quality
{
  Low,
  Medium,
  High,
  Ultra,

  Count
};

int GetStepsForQuality(Quality quality);

struct TrainInfo
{
  Quality quality_;
  int grids_per_batch_;
};

// ---- extensions -----------------------------------------------------------

// Feature vector layout fed to the decoder MLP.
//  Reference: 4 unweighted g0 corners, bilinear g1, 12 triangle waves
//             (periods 8,4,2), the normalised LOD, zero pad to a multiple of 4.
//  Bilinear:  bilinear g0, bilinear g1, 8 triangle waves (periods 4,2).
enum class FeatureLayout : int32_t
{
  Reference = 0,
  Bilinear = 1,
};

enum class Activation : int32_t
{
  HardGELU = 0,
  ReLU = 1,
};

// Fused: one gather+MLP+loss+backward+scatter kernel on BF16 MFMA with FP32
// accumulation and FP32 master state. Plain: one FP32 kernel per stage with
// every activation materialised in HBM (the reference's stage structure).
enum class TrainPath : int32_t
{
  Fused = 0,
  Plain = 1,
};

enum class EvalPrecision : int32_t
{
  FP32 = 0,
  BF16 = 1,
  FP8 = 2,
  INT8 = 3,
  BF16_MFMA = 4,  // BF16 rounding points, MFMA accumulation order (the fast BF16 decode)
  FP8_MFMA = 5,   // FP8 E4M3 rounding points and scales, MFMA accumulation order
};

struct LayoutInfo
{
  int32_t g0_bits_;
  int32_t g1_bits_;
  int32_t g0_channels_;
  int32_t g1_channels_;
  int32_t g0_scale_;  // g0 level-0 side = dim / g0_scale_ (g1 is half of g0)
  int32_t hidden_dim_;
  FeatureLayout feature_layout_;
  Activation activation_;
};

struct ContextInitInfoEx
{
  Profile profile_;        // used when use_layout_ is false
  bool use_layout_;
  LayoutInfo layout_;
  int32_t min_dimension_;  // smallest accepted texture side; 0 = kMinDimension
  int32_t mip_min_dim_;    // mip chain stops at this side; 0 = 4 (reference)
  uint32_t seed_;          // 0 = 123 (reference seed)
};

struct TrainInfoEx
{
  int32_t steps_;
  int32_t grids_per_batch_;
  float lr_grid_;     // cosine-annealed peak, reference 0.01
  float lr_decoder_;  // cosine-annealed peak, reference 0.005
  float init_lo_;     // latent init range, reference [-0.1, 0.1)
  float init_hi_;
  TrainPath path_;
};

// One in-memory texture: dim x dim RGBA8 pixels (the form the reference's
// image loader produces) plus the manifest semantics for it.
struct TextureInput
{
  const uint8_t* rgba8_;
  Semantic semantic_;
  const char* channels_;  // subset of "RGBA" in order, e.g. "RGB", "G"
  const char* name_;
};

struct EvalDetail
{
  double sse_[kMaxChannels];   // per output channel, all mips
  double count_;               // texels x mips summed (per channel)
};

class CompressedData;
struct TextureData;

class Context
{
public:
  static const int kMinDimension = 1024;
  static const int kMaxDimension = 8192;
  static const int kMaxMips = openntc::kMaxMips;
  static const int kMaxChannels = openntc::kMaxChannels;
  static const int kMaxLevels = openntc::kMaxLevels;

  Context();
  ~Context();

  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  Result Init(const ContextInitInfo& init_info);
  void Destroy();
  void BeginTraining(const TrainInfo& train_info);
  TrainProgress Train(int32_t num_batches);
  TrainProgress TrainUntilComplete();
  EvalResults Eval();

  Result GetLastError() const;
  CalibrationData Calibrate(float headroom = 1.1f);
  CompressedData GetCompressedData();
  Result LoadManifest(const std::string& filepath);
  void UnloadManifest();
  TextureData GetTextureData();
  int32_t GetMipDim(int mip) const;
  int32_t FeatureLevelForLod(int32_t mip) const;

  // extensions
  Result InitEx(const ContextInitInfoEx& init_info);
  Result LoadTextures(const TextureInput* textures, int32_t count, int32_t dim);
  void BeginTrainingEx(const TrainInfoEx& train_info);
  EvalResults EvalEx(EvalPrecision precision, EvalDetail* o_detail = nullptr);
  struct Impl;
  Impl* GetImpl() { return impl_; }

  // Host-side parameters after training

  uint32_t* g0_host_[kMaxLevels];
  uint32_t* g1_host_[kMaxLevels];
  uint32_t* g0_host_packed_ = nullptr;
  uint32_t* g1_host_packed_ = nullptr;

  uint8_t* decoder_host_;
  size_t decoder_size_;
  CalibrationData caldata_;

  uint8_t* mips_host_[kMaxSources][kMaxMips];

private:
  Impl* impl_;
};

struct TextureData
{
  int32_t tex_count_;
  int32_t mip_count_;
  int32_t channels_[kMaxSources];
  Semantic semantics_[kMaxSources];
  uint8_t* mips_[kMaxSources][Context::kMaxMips];
};

} // namespace openntc
