// GPU-free runtime half of the API: name tables, decoder constants and the
// ONTC container (libopenntc/openntc_runtime.cpp).
//
// Container layout (openntc_runtime.cpp:179-253): 16-byte header
// {0x43544E4F 'ONTC', 4, json bytes, blob bytes}, compact JSON with keys in
// sorted order, then blobs g0_0.., g1_0.., decoder, each starting at a
// multiple of 256 bytes from the end of the JSON.
#include <libopenntc/openntc_runtime.h>
#include <ntc_hip.h>

#include "json.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace openntc
{
// ---- names (openntc_runtime.cpp:20-154) --------------------------------------------
const char* SemanticToString(Semantic sem)
{
  switch (sem)
  {
  case Semantic::Albedo: return "Albedo";
  case Semantic::Alpha: return "Alpha";
  case Semantic::Displacement: return "Displacement";
  case Semantic::Emissive: return "Emissive";
  case Semantic::Gloss: return "Gloss";
  case Semantic::Metallic: return "Metallic";
  case Semantic::Normal: return "Normal";
  case Semantic::AO: return "AO";
  case Semantic::Roughness: return "Roughness";
  case Semantic::Specular: return "Specular";
  case Semantic::Transmission: return "Transmission";
  default: return "None";
  }
}

Semantic SemanticFromName(const std::string& s)
{
  struct Alias
  {
    const char* name;
    Semantic sem;
  };
  static const Alias table[] = {
      {"Albedo", Semantic::Albedo},         {"Diffuse", Semantic::Albedo},
      {"Alpha", Semantic::Alpha},           {"Mask", Semantic::Alpha},
      {"AlphaMask", Semantic::Alpha},       {"Displ", Semantic::Displacement},
      {"Displacement", Semantic::Displacement}, {"Emissive", Semantic::Emissive},
      {"Emission", Semantic::Emissive},     {"Glossiness", Semantic::Gloss},
      {"Gloss", Semantic::Gloss},           {"Metalness", Semantic::Metallic},
      {"Metallic", Semantic::Metallic},     {"Normal", Semantic::Normal},
      {"Occlusion", Semantic::AO},          {"AO", Semantic::AO},
      {"AmbientOcclusion", Semantic::AO},   {"Roughness", Semantic::Roughness},
      {"SpecularColor", Semantic::Specular}, {"Specular", Semantic::Specular},
      {"Transmission", Semantic::Transmission},
  };
  for (const Alias& a : table)
    if (s == a.name) return a.sem;
  return Semantic::None;
}

int GetChannelCountForSemantic(const Semantic sem)
{
  // None accepts any channel count (-1); Count is invalid (-1).
  static const int32_t counts[] = {-1, 3, 1, 1, 1, 1, 1, 3, 1, 1, 3, 1, -1};
  const int i = (int)sem;
  if (i < 0 || i > 12) return -1;
  return counts[i];
}

const char* ProfileToString(Profile p)
{
  switch (p)
  {
  case Profile::Bpp_0_2: return "bpp_0_2";
  case Profile::Bpp_0_5: return "bpp_0_5";
  case Profile::Bpp_1_0: return "bpp_1_0";
  case Profile::Bpp_2_25: return "bpp_2_25";
  default: return "unknown";
  }
}

bool ProfileFromString(const std::string& s, Profile& o_profile)
{
  static const Profile all[] = {Profile::Bpp_0_2, Profile::Bpp_0_5, Profile::Bpp_1_0, Profile::Bpp_2_25};
  for (Profile p : all)
    if (s == ProfileToString(p))
    {
      o_profile = p;
      return true;
    }
  return false;
}

// openntc_runtime.cpp:156-177
void FillNTCConstants(const CompressedData& data, NTCConstants& o)
{
  o = {};
  for (int i = 0; i < data.level_count_ && i < kMaxLevels; ++i)
  {
    o.g0_grid_dim_[i] = data.g0_grid_dim_[i];
    o.g1_grid_dim_[i] = data.g1_grid_dim_[i];
    o.g0_offset_[i] = (uint32_t)data.g0_offset_[i];
    o.g1_offset_[i] = (uint32_t)data.g1_offset_[i];
  }
  o.g0_bits_per_channel_ = data.g0_bits_per_channel_;
  o.g1_bits_per_channel_ = data.g1_bits_per_channel_;
  o.g0_channels_ = data.g0_channels_;
  o.g1_channels_ = data.g1_channels_;
  o.dim_ = data.dim_;
  o.mip_count_ = data.mip_count_;
  o.rcp_s_a1_ = 1.0f / data.caldata_.s_a1_;
  o.rcp_s_a2_ = 1.0f / data.caldata_.s_a2_;
  o.channel_count_ = data.channel_count_;
  for (int i = 0; i < data.channel_count_ && i < kMaxChannels; ++i)
    o.channel_semantics_[i] = (uint32_t)data.channel_semantics_[i];
}

// ---- JSON (json.h) ----------------------------------------------------------------------
namespace
{
using json::Json;
using json::Parser;

// Shortest decimal that reads back to the same double (".0" appended to
// integral values, as the reference's JSON writer does).
std::string FormatDouble(double d)
{
  char buf[64];
  for (int prec = 1; prec <= 17; ++prec)
  {
    snprintf(buf, sizeof(buf), "%.*g", prec, d);
    if (strtod(buf, nullptr) == d) break;
  }
  std::string s = buf;
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

std::string Quote(const std::string& s)
{
  std::string o = "\"";
  for (char c : s)
  {
    if (c == '"' || c == '\\')
    {
      o += '\\';
      o += c;
    }
    else
      o += c;
  }
  return o + "\"";
}

uint64_t RoundUp256(uint64_t n) { return ((n + 255) / 256) * 256; }

struct Blob
{
  std::string name;
  uint64_t size;
  const void* data;
};
} // namespace

// openntc_runtime.cpp:179-253
Result Dump(const std::string& path, const CompressedData& data)
{
  if (data.level_count_ < 1 || data.level_count_ > kMaxLevels) return Result::InvalidState;
  std::vector<Blob> blobs;
  for (int i = 0; i < data.level_count_; ++i) blobs.push_back({"g0_" + std::to_string(i), data.g0_size_[i], data.g0_[i]});
  for (int i = 0; i < data.level_count_; ++i) blobs.push_back({"g1_" + std::to_string(i), data.g1_size_[i], data.g1_[i]});
  blobs.push_back({"decoder", data.decoder_size_, data.decoder_});

  auto dims = [](const int* d) {
    std::string s = "[";
    for (int i = 0; i < kMaxLevels; ++i) s += (i ? "," : "") + std::to_string(d[i]);
    return s + "]";
  };
  std::string js = "{";
  js += "\"blobs\":[";
  uint64_t off = 0;
  for (size_t i = 0; i < blobs.size(); ++i)
  {
    off = RoundUp256(off);
    js += (i ? "," : "");
    js += "{\"name\":" + Quote(blobs[i].name) + ",\"offset\":" + std::to_string(off) +
          ",\"size\":" + std::to_string(blobs[i].size) + "}";
    off += blobs[i].size;
  }
  js += "],";
  js += "\"calibration\":{\"max_abs_a1\":" + FormatDouble(data.caldata_.max_abs_a1_) +
        ",\"max_abs_a2\":" + FormatDouble(data.caldata_.max_abs_a2_) + ",\"s_a1\":" + FormatDouble(data.caldata_.s_a1_) +
        ",\"s_a2\":" + FormatDouble(data.caldata_.s_a2_) + "},";
  js += "\"channel_semantics\":[";
  for (int i = 0; i < data.channel_count_; ++i)
    js += std::string(i ? "," : "") + Quote(SemanticToString(data.channel_semantics_[i]));
  js += "],";
  js += "\"dim\":" + std::to_string(data.dim_) + ",";
  js += "\"g0\":{\"bits\":" + std::to_string(data.g0_bits_per_channel_) + ",\"channels\":" +
        std::to_string(data.g0_channels_) + ",\"grid_dims\":" + dims(data.g0_grid_dim_) + "},";
  js += "\"g1\":{\"bits\":" + std::to_string(data.g1_bits_per_channel_) + ",\"channels\":" +
        std::to_string(data.g1_channels_) + ",\"grid_dims\":" + dims(data.g1_grid_dim_) + "},";
  js += "\"level_count\":" + std::to_string(data.level_count_) + ",";
  js += "\"mip_count\":" + std::to_string(data.mip_count_) + ",";
  js += "\"profile\":" + Quote(ProfileToString(data.profile_)) + ",";
  js += "\"source\":{\"generator\":\"openntc\",\"version\":5}";
  js += "}";

  const uint32_t header[4] = {0x43544E4Fu, 4u, (uint32_t)js.size(), (uint32_t)off};
  std::ofstream f(path, std::ios::binary);
  if (!f) return Result::FileWriteFailure;
  f.write((const char*)header, sizeof(header));
  f.write(js.data(), (std::streamsize)js.size());
  off = 0;
  for (const Blob& b : blobs)
  {
    const uint64_t aligned = RoundUp256(off);
    for (; off < aligned; ++off) f.put(0);
    if (b.size) f.write((const char*)b.data, (std::streamsize)b.size);
    off += b.size;
  }
  const bool ok = f.good();
  f.close();
  return ok ? Result::Success : Result::FileWriteFailure;
}

// openntc_runtime.cpp:307-456
Result Load(const std::string& path, FileData& o)
{
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) return Result::FileNotFound;
  const std::streamoff raw_size = f.tellg();
  if (raw_size < 16) return Result::InvalidFile;
  delete[] o.raw_;
  o.raw_ = new uint8_t[(size_t)raw_size];
  f.seekg(0);
  f.read((char*)o.raw_, raw_size);
  if (!f) return Result::InvalidFile;
  uint32_t header[4];
  memcpy(header, o.raw_, sizeof(header));
  if (header[0] != 0x43544E4Fu || header[1] != 4u) return Result::InvalidFile;
  const uint64_t json_size = header[2], blob_size = header[3];
  const uint64_t file_size = 16 + json_size + blob_size;
  const uint64_t blob_off = 16 + json_size;
  if (file_size > (uint64_t)raw_size) return Result::InvalidFile;
  Json j;
  Parser parser((const char*)o.raw_ + 16, (const char*)o.raw_ + 16 + json_size);
  if (!parser.Parse(j) || j.kind != Json::Object) return Result::InvalidFile;

  CompressedData& d = o.data_;
  d = {};
  const Json* jg0 = j.find("g0");
  const Json* jg1 = j.find("g1");
  const Json* jcal = j.find("calibration");
  if (!jg0 || !jg1 || !jcal || jg0->kind != Json::Object || jg1->kind != Json::Object ||
      jcal->kind != Json::Object)
    return Result::InvalidFile;
  auto get_int = [](const Json* obj, const char* k, int& out) {
    const Json* v = obj->find(k);
    if (!v || v->kind != Json::Number || !v->is_int) return false;
    out = (int)v->inum;
    return true;
  };
  auto get_float = [](const Json* obj, const char* k, float& out) {
    const Json* v = obj->find(k);
    if (!v || v->kind != Json::Number) return false;
    out = (float)v->num;
    return true;
  };
  auto get_dims = [](const Json* obj, const char* k, int* out) {
    const Json* v = obj->find(k);
    if (!v || v->kind != Json::Array || v->arr.size() != (size_t)kMaxLevels) return false;
    for (int i = 0; i < kMaxLevels; ++i)
    {
      if (v->arr[i].kind != Json::Number || !v->arr[i].is_int) return false;
      out[i] = (int)v->arr[i].inum;
    }
    return true;
  };
  bool ok = true;
  ok &= get_dims(jg0, "grid_dims", d.g0_grid_dim_);
  ok &= get_int(jg0, "bits", d.g0_bits_per_channel_);
  ok &= get_int(jg0, "channels", d.g0_channels_);
  ok &= get_dims(jg1, "grid_dims", d.g1_grid_dim_);
  ok &= get_int(jg1, "bits", d.g1_bits_per_channel_);
  ok &= get_int(jg1, "channels", d.g1_channels_);
  ok &= get_int(&j, "dim", d.dim_);
  ok &= get_int(&j, "mip_count", d.mip_count_);
  ok &= get_int(&j, "level_count", d.level_count_);
  ok &= get_float(jcal, "max_abs_a1", d.caldata_.max_abs_a1_);
  ok &= get_float(jcal, "s_a1", d.caldata_.s_a1_);
  ok &= get_float(jcal, "max_abs_a2", d.caldata_.max_abs_a2_);
  ok &= get_float(jcal, "s_a2", d.caldata_.s_a2_);
  if (!ok) return Result::InvalidFile;
  const Json* jp = j.find("profile");
  if (!jp || jp->kind != Json::String || !ProfileFromString(jp->str, d.profile_)) return Result::InvalidFile;
  const Json* jsem = j.find("channel_semantics");
  if (!jsem || jsem->kind != Json::Array || jsem->arr.size() > (size_t)kMaxChannels) return Result::InvalidFile;
  d.channel_count_ = (int)jsem->arr.size();
  for (int i = 0; i < d.channel_count_; ++i)
  {
    if (jsem->arr[i].kind != Json::String) return Result::InvalidFile;
    const Semantic s = SemanticFromName(jsem->arr[i].str);
    if (s == Semantic::None) return Result::InvalidFile;
    d.channel_semantics_[i] = s;
  }
  if (d.level_count_ < 1 || d.level_count_ > kMaxLevels || d.mip_count_ < 1 || d.mip_count_ > kMaxMips)
    return Result::InvalidFile;

  struct Slot
  {
    std::string name;
    void** dst;
    size_t* size;
    bool found;
  };
  std::vector<Slot> slots;
  for (int i = 0; i < d.level_count_; ++i) slots.push_back({"g0_" + std::to_string(i), (void**)&d.g0_[i], &d.g0_size_[i], false});
  for (int i = 0; i < d.level_count_; ++i) slots.push_back({"g1_" + std::to_string(i), (void**)&d.g1_[i], &d.g1_size_[i], false});
  slots.push_back({"decoder", &d.decoder_, &d.decoder_size_, false});
  const Json* jb = j.find("blobs");
  if (!jb || jb->kind != Json::Array) return Result::InvalidFile;
  for (const Json& b : jb->arr)
  {
    if (b.kind != Json::Object) return Result::InvalidFile;
    const Json* n = b.find("name");
    const Json* of = b.find("offset");
    const Json* sz = b.find("size");
    if (!n || n->kind != Json::String || !of || of->kind != Json::Number || !of->is_int || !sz ||
        sz->kind != Json::Number || !sz->is_int || of->inum < 0 || sz->inum < 0)
      return Result::InvalidFile;
    // offset and size checked one at a time against what the file holds
    // after the JSON, so no sum of two 63-bit values can wrap
    const uint64_t room = file_size - blob_off;
    if ((uint64_t)of->inum > room || (uint64_t)sz->inum > room - (uint64_t)of->inum) return Result::InvalidFile;
    for (Slot& s : slots)
      if (n->str == s.name)
      {
        *s.dst = o.raw_ + blob_off + of->inum;
        *s.size = (size_t)sz->inum;
        s.found = true;
      }
  }
  for (const Slot& s : slots)
    if (!s.found) return Result::InvalidFile;
  for (int i = 1; i < d.level_count_; ++i)
  {
    d.g0_offset_[i] = d.g0_offset_[i - 1] + d.g0_size_[i - 1];
    d.g1_offset_[i] = d.g1_offset_[i - 1] + d.g1_size_[i - 1];
  }
  const int last = d.level_count_ - 1;
  d.g0_total_size_ = d.g0_offset_[last] + d.g0_size_[last];
  d.g1_total_size_ = d.g1_offset_[last] + d.g1_size_[last];
  delete[] o.g0_packed_;
  delete[] o.g1_packed_;
  o.g0_packed_ = new uint8_t[d.g0_total_size_ ? d.g0_total_size_ : 1];
  o.g1_packed_ = new uint8_t[d.g1_total_size_ ? d.g1_total_size_ : 1];
  for (int i = 0; i < d.level_count_; ++i)
  {
    memcpy(o.g0_packed_ + d.g0_offset_[i], d.g0_[i], d.g0_size_[i]);
    memcpy(o.g1_packed_ + d.g1_offset_[i], d.g1_[i], d.g1_size_[i]);
    d.g0_[i] = (uint32_t*)(o.g0_packed_ + d.g0_offset_[i]);
    d.g1_[i] = (uint32_t*)(o.g1_packed_ + d.g1_offset_[i]);
  }
  return Result::Success;
}

FileData::FileData() : data_{}, raw_(nullptr), g0_packed_(nullptr), g1_packed_(nullptr) {}

FileData::~FileData()
{
  delete[] raw_;
  delete[] g0_packed_;
  delete[] g1_packed_;
}

const CompressedData& FileData::Data() { return data_; }
} // namespace openntc

// ---- C ABI of the container (include/ntc_hip.h) ----------------------------------------------
struct ntc_file
{
  openntc::FileData f;
};

namespace
{
void ToCompressed(const ntc_compressed& c, openntc::CompressedData& d)
{
  d = {};
  for (int i = 0; i < openntc::kMaxLevels; ++i)
  {
    d.g0_[i] = (uint32_t*)c.g0[i];
    d.g1_[i] = (uint32_t*)c.g1[i];
    d.g0_size_[i] = c.g0_size[i];
    d.g1_size_[i] = c.g1_size[i];
    d.g0_offset_[i] = c.g0_offset[i];
    d.g1_offset_[i] = c.g1_offset[i];
    d.g0_grid_dim_[i] = c.g0_grid_dim[i];
    d.g1_grid_dim_[i] = c.g1_grid_dim[i];
  }
  d.decoder_ = (void*)c.decoder;
  d.g0_total_size_ = c.g0_total_size;
  d.g1_total_size_ = c.g1_total_size;
  d.decoder_size_ = c.decoder_size;
  d.g0_bits_per_channel_ = c.g0_bits;
  d.g1_bits_per_channel_ = c.g1_bits;
  d.g0_channels_ = c.g0_channels;
  d.g1_channels_ = c.g1_channels;
  d.dim_ = c.dim;
  d.mip_count_ = c.mip_count;
  d.level_count_ = c.level_count;
  d.channel_count_ = c.channel_count;
  for (int i = 0; i < openntc::kMaxChannels; ++i) d.channel_semantics_[i] = (openntc::Semantic)c.channel_semantics[i];
  d.profile_ = (openntc::Profile)c.profile;
  d.caldata_ = {c.max_abs_a1, c.max_abs_a2, c.s_a1, c.s_a2};
}

void FromCompressed(const openntc::CompressedData& d, ntc_compressed& c)
{
  c = {};
  for (int i = 0; i < openntc::kMaxLevels; ++i)
  {
    c.g0[i] = d.g0_[i];
    c.g1[i] = d.g1_[i];
    c.g0_size[i] = d.g0_size_[i];
    c.g1_size[i] = d.g1_size_[i];
    c.g0_offset[i] = d.g0_offset_[i];
    c.g1_offset[i] = d.g1_offset_[i];
    c.g0_grid_dim[i] = d.g0_grid_dim_[i];
    c.g1_grid_dim[i] = d.g1_grid_dim_[i];
  }
  c.decoder = d.decoder_;
  c.g0_total_size = d.g0_total_size_;
  c.g1_total_size = d.g1_total_size_;
  c.decoder_size = d.decoder_size_;
  c.g0_bits = d.g0_bits_per_channel_;
  c.g1_bits = d.g1_bits_per_channel_;
  c.g0_channels = d.g0_channels_;
  c.g1_channels = d.g1_channels_;
  c.dim = d.dim_;
  c.mip_count = d.mip_count_;
  c.level_count = d.level_count_;
  c.channel_count = d.channel_count_;
  for (int i = 0; i < openntc::kMaxChannels; ++i) c.channel_semantics[i] = (int32_t)d.channel_semantics_[i];
  c.profile = (int32_t)d.profile_;
  c.max_abs_a1 = d.caldata_.max_abs_a1_;
  c.max_abs_a2 = d.caldata_.max_abs_a2_;
  c.s_a1 = d.caldata_.s_a1_;
  c.s_a2 = d.caldata_.s_a2_;
}
} // namespace

extern "C" {
int ntc_container_dump(const char* path, const ntc_compressed* data)
{
  if (!path || !data) return (int)openntc::Result::InvalidState;
  openntc::CompressedData d;
  ToCompressed(*data, d);
  try
  {
    return (int)openntc::Dump(path, d);
  }
  catch (const std::bad_alloc&)
  {
    return (int)openntc::Result::AllocationFailure;
  }
}

int ntc_container_load(const char* path, ntc_file** out)
{
  if (!out) return (int)openntc::Result::InvalidState;
  *out = nullptr;
  if (!path) return (int)openntc::Result::FileNotFound;
  try
  {
    std::unique_ptr<ntc_file> f(new ntc_file());
    const openntc::Result r = openntc::Load(path, f->f);
    if (r != openntc::Result::Success) return (int)r;
    *out = f.release();
    return 0;
  }
  catch (const std::bad_alloc&)
  {
    return (int)openntc::Result::AllocationFailure;
  }
}

int ntc_file_data(const ntc_file* f, ntc_compressed* out)
{
  if (!f || !out) return (int)openntc::Result::InvalidState;
  FromCompressed(const_cast<ntc_file*>(f)->f.Data(), *out);
  return 0;
}

void ntc_file_free(ntc_file* f) { delete f; }

int ntc_fill_constants(const ntc_compressed* data, void* constants240)
{
  if (!data || !constants240) return (int)openntc::Result::InvalidState;
  openntc::CompressedData d;
  ToCompressed(*data, d);
  openntc::NTCConstants k;
  openntc::FillNTCConstants(d, k);
  std::memcpy(constants240, &k, sizeof(k));
  return 0;
}
} // extern "C"

