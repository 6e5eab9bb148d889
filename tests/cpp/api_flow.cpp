// A C++ caller of the library through the public openntc::Context API, the
// way the reference's README drives it (libopenntc/libopenntc.h:8-332): the
// Tensor wrappers (libopenntc.cpp:34-269), the manifest -> train -> eval ->
// calibrate -> export -> container flow with the housekeeping calls
// (GetTextureData :1425, GetMipDim :1495, UnloadManifest :1369, the Train(n)
// clamp and the TrainComplete result once no steps remain :550-562), and the
// CudaFailure / AllocationFailure paths (:17-30, :506-527) reached through the
// library's fault-injection switch.
//
// usage: api_flow <manifest.json> <output dir> <albedo.raw>
//   albedo.raw: the manifest's Albedo texture as dim x dim x 3 bytes.
// Prints "api_flow: N checks passed" and exits 0, or names the first failed
// check and exits 1.
#include <libopenntc/libopenntc.h>
#include <ntc_hip.h>

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

using namespace openntc;

static int g_checks = 0;

#define CHECK(cond)                                                              \
  do                                                                             \
  {                                                                              \
    ++g_checks;                                                                  \
    if (!(cond))                                                                 \
    {                                                                            \
      std::fprintf(stderr, "api_flow: check failed at line %d: %s\n", __LINE__, #cond); \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

// The library's uniform draw: lo + (hi - lo) * u, u = 24 high bits of one
// mt19937 output / 2^24, in float.
static float draw(std::mt19937& g, float lo, float hi)
{
  const float u = (float)(g() >> 8) * 0x1p-24f;
  return lo + (hi - lo) * u;
}

template <typename T>
static std::vector<T> download(const T* dev, size_t n)
{
  std::vector<T> h(n);
  CHECK(hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess);
  return h;
}

static void tensors()
{
  // Tensor2d: shape, zero fill, Kaiming-uniform fill (bound 1/sqrt(columns))
  Tensor2d t;
  CHECK(t.Init(3, 5) == Result::Success);
  CHECK(t.NumElems() == 15 && t.SizeBytes() == 60 && t.DevicePtr() != nullptr);
  t.FillZero();
  for (float v : download(t.DevicePtr(), 15)) CHECK(v == 0.0f);
  std::mt19937 g(7), g2(7);
  t.FillKaiming(g);
  const float b = 1.0f / std::sqrt(5.0f);
  const std::vector<float> k = download(t.DevicePtr(), 15);
  for (int i = 0; i < 15; ++i)
  {
    CHECK(k[i] == draw(g2, -b, b));
    CHECK(k[i] >= -b && k[i] < b);
  }
  Tensor2d like;
  CHECK(like.InitLike(t) == Result::Success && like.NumElems() == 15);
  t.Destroy();
  CHECK(t.DevicePtr() == nullptr);
  t.Destroy();  // a second Destroy is harmless
  CHECK(t.Init(2, 2) == Result::Success && t.NumElems() == 4);

  // Tensor3d
  Tensor3d t3;
  CHECK(!t3.IsInitialized());
  CHECK(t3.Init(2, 3, 4) == Result::Success && t3.IsInitialized());
  CHECK(t3.NumElems() == 24 && t3.SizeBytes() == 96);
  std::mt19937 h(11), h2(11);
  t3.FillUniform(h, -0.25f, 0.75f);
  const std::vector<float> u = download(t3.DevicePtr(), 24);
  for (int i = 0; i < 24; ++i) CHECK(u[i] == draw(h2, -0.25f, 0.75f));
  t3.FillZero();
  for (float v : download(t3.DevicePtr(), 24)) CHECK(v == 0.0f);
  Tensor3d l3;
  CHECK(l3.InitLike(t3) == Result::Success && l3.NumElems() == 24);
  t3.Destroy();
  CHECK(!t3.IsInitialized());

  // U8Tensor3d
  U8Tensor3d u8;
  CHECK(u8.Init(4, 4, 3) == Result::Success && u8.IsInitialized());
  CHECK(u8.NumElems() == 48 && u8.SizeBytes() == 48);
  std::vector<uint8_t> ab(48, 0xAB);
  CHECK(hipMemcpy(u8.DevicePtr(), ab.data(), 48, hipMemcpyHostToDevice) == hipSuccess);
  u8.FillZero();
  for (uint8_t v : download(u8.DevicePtr(), 48)) CHECK(v == 0);
  U8Tensor3d l8;
  CHECK(l8.InitLike(u8) == Result::Success && l8.NumElems() == 48);
  u8.Destroy();
  CHECK(!u8.IsInitialized());

  // IntTensor1d
  IntTensor1d it;
  CHECK(it.Init(10) == Result::Success && it.DevicePtr() != nullptr);
  std::vector<int> iv(10);
  for (int i = 0; i < 10; ++i) iv[i] = i * i - 7;
  CHECK(hipMemcpy(it.DevicePtr(), iv.data(), 40, hipMemcpyHostToDevice) == hipSuccess);
  CHECK(download(it.DevicePtr(), 10) == iv);
  it.Destroy();
  CHECK(it.DevicePtr() == nullptr);

  // a failed allocation is AllocationFailure; the next one succeeds
  CHECK(ntc_debug_inject_fault(NTC_FAULT_ALLOC, 1) == 0);
  Tensor2d f;
  CHECK(f.Init(4, 4) == Result::AllocationFailure);
  CHECK(f.Init(4, 4) == Result::Success);
  CHECK(ntc_debug_inject_fault(NTC_FAULT_ALLOC, 2) == 0);
  Tensor3d f3;
  IntTensor1d fi;
  CHECK(f3.Init(2, 2, 2) == Result::Success);
  CHECK(fi.Init(3) == Result::AllocationFailure);
  CHECK(ntc_debug_inject_fault(NTC_FAULT_ALLOC, 1) == 0 && ntc_debug_inject_fault(NTC_FAULT_ALLOC, 0) == 0);
  CHECK(fi.Init(3) == Result::Success);  // disarmed
}

static std::vector<uint8_t> read_file(const std::string& p)
{
  std::ifstream f(p, std::ios::binary);
  CHECK(f.good());
  return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

static void flow(const std::string& manifest, const std::string& out_dir, const std::string& raw_path)
{
  Context ctx;
  CHECK(ctx.GetMipDim(0) == 0);
  // a refused Train reports InvalidState in its progress and leaves the last error
  CHECK(ctx.Train(1).result_ == Result::InvalidState && ctx.GetLastError() == Result::Success);
  CHECK(ctx.Init(ContextInitInfo{Profile::Bpp_0_5}) == Result::Success);
  CHECK(ctx.GetTextureData().tex_count_ == 0);
  CHECK(ctx.GetMipDim(0) == 0);
  CHECK(ctx.LoadManifest(manifest) == Result::Success);

  // mip sides: 1024 halving down to 4
  const int dim = ctx.GetMipDim(0);
  CHECK(dim == 1024);
  int mips = 0;
  while (ctx.GetMipDim(mips) > 0) ++mips;
  CHECK(mips == 9 && ctx.GetMipDim(8) == 4 && ctx.GetMipDim(-1) == 0 && ctx.GetMipDim(99) == 0);
  for (int m = 1; m < mips; ++m) CHECK(ctx.GetMipDim(m) == ctx.GetMipDim(m - 1) / 2);

  // host copies of the sources' mips: one source per semantic, in semantic
  // order (Albedo RGB, Metallic, Normal RGB, AO, Roughness)
  const TextureData td = ctx.GetTextureData();
  CHECK(td.tex_count_ == 5 && td.mip_count_ == mips);
  const Semantic order[5] = {Semantic::Albedo, Semantic::Metallic, Semantic::Normal, Semantic::AO,
                             Semantic::Roughness};
  const int chans[5] = {3, 1, 3, 1, 1};
  for (int i = 0; i < 5; ++i) CHECK(td.semantics_[i] == order[i] && td.channels_[i] == chans[i]);
  for (int i = 0; i < td.tex_count_; ++i)
    for (int m = 0; m < mips; ++m) CHECK(td.mips_[i][m] != nullptr);
  const std::vector<uint8_t> raw = read_file(raw_path);
  CHECK(raw.size() == (size_t)dim * dim * 3);
  CHECK(std::memcmp(td.mips_[0][0], raw.data(), raw.size()) == 0);

  // training: the reference's quality preset, Train(n) clamped to what is left
  ctx.BeginTraining(TrainInfo{Quality::Low, 1});
  const int total = GetStepsForQuality(Quality::Low);
  TrainProgress p = ctx.Train(100);
  CHECK(p.phase_ == TrainPhase::TrainInProgress && p.result_ == Result::Success);
  CHECK(p.batches_complete_ == 100 && p.total_batches_ == total);
  p = ctx.Train(1 << 30);
  CHECK(p.phase_ == TrainPhase::TrainComplete && p.result_ == Result::Success);
  CHECK(p.batches_complete_ == total && p.total_batches_ == total);
  p = ctx.Train(5);  // nothing left: TrainComplete, no step taken
  CHECK(p.phase_ == TrainPhase::TrainComplete && p.result_ == Result::Success && p.batches_complete_ == total);
  p = ctx.TrainUntilComplete();
  CHECK(p.phase_ == TrainPhase::TrainComplete && p.batches_complete_ == total);

  const EvalResults ev = ctx.Eval();
  CHECK(std::isfinite(ev.psnr_) && ev.psnr_ > 15.0 && ev.mse_ > 0.0);
  std::printf("api_flow: %d steps, PSNR %.3f dB\n", total, ev.psnr_);
  const CalibrationData cal = ctx.Calibrate();
  CHECK(cal.max_abs_a1_ > 0.0f && cal.max_abs_a2_ > 0.0f);
  CHECK(ctx.GetLastError() == Result::Success);

  // export and the ONTC container
  const CompressedData cd = ctx.GetCompressedData();
  CHECK(cd.dim_ == dim && cd.mip_count_ == mips && cd.level_count_ > 0 && cd.channel_count_ == 9);
  CHECK(cd.profile_ == Profile::Bpp_0_5 && cd.decoder_size_ > 0 && cd.g0_total_size_ > 0);
  const std::string path = out_dir + "/set.ontc";
  CHECK(Dump(path, cd) == Result::Success);
  FileData fd;
  CHECK(Load(path, fd) == Result::Success);
  const CompressedData& ld = fd.Data();
  CHECK(ld.dim_ == cd.dim_ && ld.mip_count_ == cd.mip_count_ && ld.level_count_ == cd.level_count_);
  CHECK(ld.g0_total_size_ == cd.g0_total_size_ && ld.g1_total_size_ == cd.g1_total_size_);
  CHECK(ld.decoder_size_ == cd.decoder_size_);
  for (int l = 0; l < cd.level_count_; ++l)
  {
    CHECK(ld.g0_size_[l] == cd.g0_size_[l] && std::memcmp(ld.g0_[l], cd.g0_[l], cd.g0_size_[l]) == 0);
    CHECK(ld.g1_size_[l] == cd.g1_size_[l] && std::memcmp(ld.g1_[l], cd.g1_[l], cd.g1_size_[l]) == 0);
  }
  CHECK(std::memcmp(ld.decoder_, cd.decoder_, cd.decoder_size_) == 0);
  CHECK(std::memcmp(&ld.caldata_, &cd.caldata_, sizeof(CalibrationData)) == 0);
  NTCConstants a = {}, b = {};
  FillNTCConstants(cd, a);
  FillNTCConstants(ld, b);
  CHECK(std::memcmp(&a, &b, sizeof(a)) == 0);
  CHECK(a.dim_ == dim && a.mip_count_ == mips && a.channel_count_ == 9);

  // unload: no texture data, no mips, training refused; a reload works
  ctx.UnloadManifest();
  CHECK(ctx.GetTextureData().tex_count_ == 0 && ctx.GetMipDim(0) == 0);
  CHECK(ctx.Train(1).result_ == Result::InvalidState);
  ctx.UnloadManifest();  // a second unload is harmless
  CHECK(ctx.LoadManifest(manifest) == Result::Success && ctx.GetMipDim(0) == dim);
  CHECK(ctx.GetTextureData().tex_count_ == 5);
  ctx.Destroy();
}

static void faults(const std::string& manifest)
{
  Context c;
  CHECK(c.Init(ContextInitInfo{Profile::Bpp_0_5}) == Result::Success);
  // an allocation failing inside LoadManifest: AllocationFailure, nothing loaded
  CHECK(ntc_debug_inject_fault(NTC_FAULT_ALLOC, 1) == 0);
  CHECK(c.LoadManifest(manifest) == Result::AllocationFailure);
  CHECK(c.GetTextureData().tex_count_ == 0 && c.GetMipDim(0) == 0);
  CHECK(c.Train(1).result_ == Result::InvalidState);
  // a copy failing inside LoadManifest: CudaFailure
  CHECK(ntc_debug_inject_fault(NTC_FAULT_COPY, 1) == 0);
  CHECK(c.LoadManifest(manifest) == Result::CudaFailure);
  CHECK(c.GetTextureData().tex_count_ == 0);
  CHECK(c.LoadManifest(manifest) == Result::Success);

  // an allocation failing inside BeginTraining: the error is recorded and
  // training is refused until the manifest is loaded again
  CHECK(ntc_debug_inject_fault(NTC_FAULT_ALLOC, 1) == 0);
  c.BeginTraining(TrainInfo{Quality::Low, 1});
  CHECK(c.GetLastError() == Result::AllocationFailure);
  TrainProgress p = c.Train(1);
  CHECK(p.phase_ == TrainPhase::TrainError && p.result_ == Result::InvalidState);
  CHECK(c.LoadManifest(manifest) == Result::Success);
  c.BeginTraining(TrainInfo{Quality::Low, 1});
  p = c.Train(10);
  CHECK(p.phase_ == TrainPhase::TrainInProgress && p.batches_complete_ == 10);

  // a launch failing inside Train: TrainError / CudaFailure at the step it hit
  CHECK(ntc_debug_inject_fault(NTC_FAULT_LAUNCH, 1) == 0);
  p = c.Train(10);
  CHECK(p.phase_ == TrainPhase::TrainError && p.result_ == Result::CudaFailure);
  CHECK(p.batches_complete_ == 10 && c.GetLastError() == Result::CudaFailure);
  // a copy failing inside Eval: the error is recorded, the results are empty
  CHECK(c.LoadManifest(manifest) == Result::Success);
  c.BeginTraining(TrainInfo{Quality::Low, 1});
  CHECK(c.Train(5).result_ == Result::Success);
  CHECK(ntc_debug_inject_fault(NTC_FAULT_COPY, 1) == 0);
  (void)c.Eval();
  CHECK(c.GetLastError() == Result::CudaFailure);
  CHECK(ntc_debug_inject_fault(NTC_FAULT_LAUNCH, 0) == 0 && ntc_debug_inject_fault(NTC_FAULT_COPY, 0) == 0);
  CHECK(ntc_debug_inject_fault(NTC_FAULT_ALLOC, 0) == 0);
  CHECK(std::isfinite(c.Eval().psnr_));
  // unknown sites and negative counts are refused
  CHECK(ntc_debug_inject_fault(0, 1) == NTC_E_INVALID && ntc_debug_inject_fault(4, 1) == NTC_E_INVALID);
  CHECK(ntc_debug_inject_fault(NTC_FAULT_COPY, -1) == NTC_E_INVALID);
}

int main(int argc, char** argv)
{
  if (argc != 4)
  {
    std::fprintf(stderr, "usage: api_flow <manifest.json> <output dir> <albedo.raw>\n");
    return 2;
  }
  tensors();
  std::printf("api_flow: tensors done\n");
  std::fflush(stdout);
  flow(argv[1], argv[2], argv[3]);
  std::printf("api_flow: flow done\n");
  std::fflush(stdout);
  faults(argv[1]);
  std::printf("api_flow: %d checks passed\n", g_checks);
  return 0;
}
