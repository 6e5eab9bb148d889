// Texture manifest front end of Context::LoadManifest (libopenntc.cpp:
// 1103-1207 validation, 1296-1313 image loading): the JSON manifest is
// checked with the reference's rules and error codes, every source image is
// decoded to RGBA8, and the texture set is built by the same code as
// Context::LoadTextures (semantic sort, mips, packages).
//
// The reference decodes images with a vendored third-party loader; this file
// has its own decoders for the formats texture manifests use: PNG (every
// colour type and bit depth, tRNS, Adam7 interlacing; inflate from the system
// zlib), TGA (true colour and grey, uncompressed or RLE) and binary PNM
// (P5/P6). Images become 4 channels the way the reference requests them:
// grey g -> (g, g, g, 255), grey+alpha -> (g, g, g, a), RGB -> (r, g, b, 255);
// 16-bit samples keep their high byte.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "context_impl.h"
#include "images.h"
#include "json.h"

using namespace openntc;

namespace
{
bool ReadFile(const std::string& path, std::vector<uint8_t>& out)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.seekg(0, std::ios::end);
  const std::streamoff n = f.tellg();
  if (n < 0) return false;
  f.seekg(0, std::ios::beg);
  out.resize((size_t)n);
  return n == 0 || (bool)f.read(reinterpret_cast<char*>(out.data()), n);
}

uint32_t Be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// ---- PNG (ISO/IEC 15948) -----------------------------------------------------------
int Paeth(int a, int b, int c)
{
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

// Undoes the scanline filters of one (sub)image of w x h pixels in place;
// bpp: bytes per complete pixel (>= 1), stride: bytes per row.
bool Unfilter(uint8_t* data, size_t avail, int w, int h, int bpp, size_t stride, std::vector<uint8_t>& out)
{
  out.assign(stride * (size_t)h, 0);
  if (w == 0 || h == 0) return true;
  if (avail < (stride + 1) * (size_t)h) return false;
  for (int y = 0; y < h; ++y)
  {
    const uint8_t ft = data[(stride + 1) * y];
    const uint8_t* src = data + (stride + 1) * y + 1;
    uint8_t* row = out.data() + stride * y;
    const uint8_t* up = y ? row - stride : nullptr;
    for (size_t i = 0; i < stride; ++i)
    {
      const int a = i >= (size_t)bpp ? row[i - bpp] : 0, b = up ? up[i] : 0;
      const int c = (up && i >= (size_t)bpp) ? up[i - bpp] : 0;
      int v = src[i];
      switch (ft)
      {
      case 0: break;
      case 1: v += a; break;
      case 2: v += b; break;
      case 3: v += (a + b) >> 1; break;
      case 4: v += Paeth(a, b, c); break;
      default: return false;
      }
      row[i] = (uint8_t)v;
    }
  }
  return true;
}

struct PngInfo
{
  int w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> plte, trns;
};

int Channels(int ctype) { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }

// Sample k of row `row` (unfiltered bytes) at bit depth d.
int Sample(const uint8_t* row, size_t k, int d)
{
  if (d == 8) return row[k];
  if (d == 16) return (row[2 * k] << 8) | row[2 * k + 1];
  const size_t bit = k * (size_t)d;
  return (row[bit >> 3] >> (8 - d - (int)(bit & 7))) & ((1 << d) - 1);
}

// Writes the pixels of an unfiltered (sub)image into rgba at the given
// positions of the full image.
void Expand(const PngInfo& p, const uint8_t* rows, size_t stride, int w, int h, int x0, int y0, int dx, int dy,
            uint8_t* rgba)
{
  const int nc = Channels(p.ctype), d = p.depth;
  const int maxv = (1 << d) - 1;
  auto to8 = [&](int v) -> uint8_t { return d == 16 ? (uint8_t)(v >> 8) : (uint8_t)(d == 8 ? v : v * 255 / maxv); };
  for (int y = 0; y < h; ++y)
  {
    const uint8_t* row = rows + stride * y;
    for (int x = 0; x < w; ++x)
    {
      uint8_t* o = rgba + 4 * ((size_t)(y0 + y * dy) * p.w + (x0 + x * dx));
      if (p.ctype == 3)
      {
        const int idx = Sample(row, x, d);
        const bool ok = (size_t)(3 * idx + 2) < p.plte.size();
        o[0] = ok ? p.plte[3 * idx] : 0;
        o[1] = ok ? p.plte[3 * idx + 1] : 0;
        o[2] = ok ? p.plte[3 * idx + 2] : 0;
        o[3] = (size_t)idx < p.trns.size() ? p.trns[idx] : 255;
        continue;
      }
      int s[4] = {0, 0, 0, 0};
      for (int c = 0; c < nc; ++c) s[c] = Sample(row, (size_t)x * nc + c, d);
      if (p.ctype == 0 || p.ctype == 4)
      {
        o[0] = o[1] = o[2] = to8(s[0]);
        o[3] = p.ctype == 4 ? to8(s[1]) : 255;
        if (p.ctype == 0 && p.trns.size() >= 2 && s[0] == ((p.trns[0] << 8) | p.trns[1])) o[3] = 0;
      }
      else
      {
        o[0] = to8(s[0]);
        o[1] = to8(s[1]);
        o[2] = to8(s[2]);
        o[3] = p.ctype == 6 ? to8(s[3]) : 255;
        if (p.ctype == 2 && p.trns.size() >= 6 && s[0] == ((p.trns[0] << 8) | p.trns[1]) &&
            s[1] == ((p.trns[2] << 8) | p.trns[3]) && s[2] == ((p.trns[4] << 8) | p.trns[5]))
          o[3] = 0;
      }
    }
  }
}

bool DecodePng(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h)
{
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  if (f.size() < 8 || memcmp(f.data(), sig, 8) != 0) return false;
  PngInfo p;
  std::vector<uint8_t> z;
  size_t pos = 8;
  bool ihdr = false, iend = false;
  while (pos + 12 <= f.size() && !iend)
  {
    const uint32_t len = Be32(&f[pos]);
    if (len > f.size() - pos - 12) return false;
    const uint8_t* type = &f[pos + 4];
    const uint8_t* d = &f[pos + 8];
    if (!memcmp(type, "IHDR", 4))
    {
      if (len < 13) return false;
      p.w = (int)Be32(d);
      p.h = (int)Be32(d + 4);
      p.depth = d[8];
      p.ctype = d[9];
      p.interlace = d[12];
      if (d[10] != 0 || d[11] != 0 || p.interlace > 1) return false;
      ihdr = true;
    }
    else if (!memcmp(type, "PLTE", 4))
      p.plte.assign(d, d + len);
    else if (!memcmp(type, "tRNS", 4))
      p.trns.assign(d, d + len);
    else if (!memcmp(type, "IDAT", 4))
      z.insert(z.end(), d, d + len);
    else if (!memcmp(type, "IEND", 4))
      iend = true;
    pos += 12 + len;
  }
  if (!ihdr || !img::SaneImageSize(p.w, p.h)) return false;
  const int d = p.depth;
  const bool depth_ok = (p.ctype == 0 && (d == 1 || d == 2 || d == 4 || d == 8 || d == 16)) ||
                        (p.ctype == 3 && (d == 1 || d == 2 || d == 4 || d == 8)) ||
                        ((p.ctype == 2 || p.ctype == 4 || p.ctype == 6) && (d == 8 || d == 16));
  if (!depth_ok || (p.ctype == 3 && p.plte.empty())) return false;
  const int nc = Channels(p.ctype);
  const int bpp = std::max(1, nc * d / 8);
  auto stride_of = [&](int ww) { return ((size_t)ww * nc * d + 7) / 8; };
  // raw size of all passes
  static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
  static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
  size_t raw_size = 0;
  if (p.interlace == 0)
    raw_size = (stride_of(p.w) + 1) * (size_t)p.h;
  else
    for (int k = 0; k < 7; ++k)
    {
      const int pw = (p.w - ax0[k] + adx[k] - 1) / adx[k], ph = (p.h - ay0[k] + ady[k] - 1) / ady[k];
      if (pw > 0 && ph > 0) raw_size += (stride_of(pw) + 1) * (size_t)ph;
    }
  // deflate expands at most 1032:1 (RFC 1951 stored / fixed-code bounds):
  // data claiming more cannot be in the file
  if (raw_size > 1032 * (uint64_t)z.size() + 64) return false;
  std::vector<uint8_t> raw(raw_size);
  z_stream zs = {};
  if (inflateInit(&zs) != Z_OK) return false;
  zs.next_in = z.data();
  zs.avail_in = (uInt)z.size();
  zs.next_out = raw.data();
  zs.avail_out = (uInt)raw.size();
  const int zr = inflate(&zs, Z_FINISH);
  const size_t got = raw.size() - zs.avail_out;
  inflateEnd(&zs);
  if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR && zr != Z_OK) || got != raw.size()) return false;
  rgba.assign((size_t)p.w * p.h * 4, 0);
  std::vector<uint8_t> rows;
  if (p.interlace == 0)
  {
    if (!Unfilter(raw.data(), raw.size(), p.w, p.h, bpp, stride_of(p.w), rows)) return false;
    Expand(p, rows.data(), stride_of(p.w), p.w, p.h, 0, 0, 1, 1, rgba.data());
  }
  else
  {
    size_t off = 0;
    for (int k = 0; k < 7; ++k)
    {
      const int pw = (p.w - ax0[k] + adx[k] - 1) / adx[k], ph = (p.h - ay0[k] + ady[k] - 1) / ady[k];
      if (pw <= 0 || ph <= 0) continue;
      const size_t st = stride_of(pw);
      if (!Unfilter(raw.data() + off, raw.size() - off, pw, ph, bpp, st, rows)) return false;
      Expand(p, rows.data(), st, pw, ph, ax0[k], ay0[k], adx[k], ady[k], rgba.data());
      off += (st + 1) * (size_t)ph;
    }
  }
  w = p.w;
  h = p.h;
  return true;
}

// ---- TGA: true colour (types 2, 10) and grey (3, 11), 8/24/32 bits ------------------
bool DecodeTga(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h)
{
  if (f.size() < 18) return false;
  const int idlen = f[0], cmtype = f[1], type = f[2], bits = f[16], desc = f[17];
  const int cmlen = f[5] | (f[6] << 8), cmbits = f[7];
  w = f[12] | (f[13] << 8);
  h = f[14] | (f[15] << 8);
  const bool rle = type == 10 || type == 11, grey = type == 3 || type == 11;
  if (cmtype > 1 || !(type == 2 || type == 3 || type == 10 || type == 11) || w <= 0 || h <= 0) return false;
  if (grey ? bits != 8 : (bits != 24 && bits != 32)) return false;
  if (!img::SaneImageSize(w, h)) return false;
  const int bypp = bits / 8;
  size_t pos = 18 + (size_t)idlen + (cmtype ? (size_t)cmlen * ((cmbits + 7) / 8) : 0);
  if (pos > f.size()) return false;
  const size_t n = (size_t)w * h;
  // the file must hold the pixels (raw) or their packets (RLE: one header
  // byte per at most 128 pixels) before anything is sized from the header
  if (!rle ? f.size() - pos < n * bypp : n > 128 * (f.size() - pos)) return false;
  std::vector<uint8_t> px(n * bypp);
  if (!rle)
    memcpy(px.data(), &f[pos], px.size());
  else
  {
    size_t k = 0;
    while (k < n)
    {
      if (pos >= f.size()) return false;
      const int hdr = f[pos++];
      const size_t cnt = (size_t)(hdr & 0x7F) + 1;
      if (k + cnt > n) return false;
      if (hdr & 0x80)
      {
        if (pos + bypp > f.size()) return false;
        for (size_t i = 0; i < cnt; ++i) memcpy(&px[(k + i) * bypp], &f[pos], bypp);
        pos += bypp;
      }
      else
      {
        if (pos + cnt * bypp > f.size()) return false;
        memcpy(&px[k * bypp], &f[pos], cnt * bypp);
        pos += cnt * bypp;
      }
      k += cnt;
    }
  }
  const bool top_left = (desc & 0x20) != 0;
  rgba.assign(n * 4, 0);
  for (int y = 0; y < h; ++y)
  {
    const int sy = top_left ? y : h - 1 - y;
    for (int x = 0; x < w; ++x)
    {
      const uint8_t* s = &px[((size_t)sy * w + x) * bypp];
      uint8_t* o = &rgba[((size_t)y * w + x) * 4];
      if (grey)
        o[0] = o[1] = o[2] = s[0], o[3] = 255;
      else
        o[0] = s[2], o[1] = s[1], o[2] = s[0], o[3] = bypp == 4 ? s[3] : 255;
    }
  }
  return true;
}

// ---- binary PNM: P5 (grey) and P6 (RGB), maxval up to 65535 ------------------------
bool DecodePnm(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h)
{
  if (f.size() < 3 || f[0] != 'P' || (f[1] != '5' && f[1] != '6')) return false;
  const int nc = f[1] == '5' ? 1 : 3;
  size_t pos = 2;
  long vals[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k)
  {
    for (;;)
    {
      while (pos < f.size() && isspace(f[pos])) ++pos;
      if (pos < f.size() && f[pos] == '#')
      {
        while (pos < f.size() && f[pos] != '\n') ++pos;
        continue;
      }
      break;
    }
    if (pos >= f.size() || !isdigit(f[pos])) return false;
    while (pos < f.size() && isdigit(f[pos]))
    {
      vals[k] = vals[k] * 10 + (f[pos++] - '0');
      if (vals[k] > 1 << 20) return false;
    }
  }
  if (pos >= f.size() || !isspace(f[pos])) return false;
  ++pos;
  w = (int)vals[0];
  h = (int)vals[1];
  const long maxv = vals[2];
  if (!img::SaneImageSize(w, h) || maxv <= 0 || maxv > 65535) return false;
  const int bps = maxv > 255 ? 2 : 1;
  const size_t n = (size_t)w * h;
  if (f.size() < pos + n * nc * bps) return false;
  rgba.assign(n * 4, 0);
  for (size_t i = 0; i < n; ++i)
  {
    uint8_t v[3];
    for (int c = 0; c < nc; ++c)
    {
      const uint8_t* s = &f[pos + (i * nc + c) * bps];
      const long x = bps == 2 ? (s[0] << 8 | s[1]) : s[0];
      v[c] = (uint8_t)(maxv == 255 ? x : (maxv == 65535 ? x >> 8 : (x * 255 + maxv / 2) / maxv));
    }
    uint8_t* o = &rgba[i * 4];
    o[0] = v[0];
    o[1] = nc == 3 ? v[1] : v[0];
    o[2] = nc == 3 ? v[2] : v[0];
    o[3] = 255;
  }
  return true;
}

bool DecodeImage(const std::string& path, std::vector<uint8_t>& rgba, int& w, int& h)
{
  // an allocation that fails fails the load (FileNotFound, as a null
  // stbi_load does in the reference), never the process
  try
  {
    std::vector<uint8_t> f;
    if (!ReadFile(path, f)) return false;
    // the order stb_image tries formats in (those read here)
    if (img::DecodeJpeg(f, rgba, w, h)) return true;
    if (DecodePng(f, rgba, w, h)) return true;
    if (img::DecodeBmp(f, rgba, w, h)) return true;
    if (DecodePnm(f, rgba, w, h)) return true;
    return DecodeTga(f, rgba, w, h);
  }
  catch (const std::bad_alloc&)
  {
    rgba.clear();
    return false;
  }
}

// libopenntc.cpp:1103-1121
bool ValidSemanticChannels(Semantic sem, const std::string& channels)
{
  static const std::string rgba = "RGBA";
  if (channels.empty() || channels.size() > 4) return false;
  size_t p = 0;
  for (char c : channels)
  {
    while (p < 4 && rgba[p] != c) ++p;
    if (p >= 4) return false;
    ++p;
  }
  return sem == Semantic::None || GetChannelCountForSemantic(sem) == (int)channels.size();
}

struct ManifestEntry
{
  std::string path, name, channels;
  Semantic semantic = Semantic::None;
  bool srgb = false, flip = false;
};

// The manifest checks of libopenntc.cpp:1125-1207, in the reference's order:
// sources in texture order, within a texture in key order of its semantics
// object (sorted), then stably sorted by semantic.
Result ParseManifest(const std::string& filepath, int min_dim, int& dim, std::vector<ManifestEntry>& out)
{
  out.clear();
  std::vector<uint8_t> text;
  if (!ReadFile(filepath, text)) return Result::FileNotFound;
  json::Json j;
  json::Parser parser(reinterpret_cast<const char*>(text.data()), reinterpret_cast<const char*>(text.data()) + text.size());
  if (!parser.Parse(j) || j.kind != json::Json::Object) return Result::InvalidManifest;
  const json::Json* tex = j.find("textures");
  if (!tex) return Result::InvalidManifest;
  auto int_or = [&](const char* k) -> long long {
    const json::Json* v = j.find(k);
    return v && v->kind == json::Json::Number ? (v->is_int ? v->inum : (long long)v->num) : -1;
  };
  const long long width = int_or("width"), height = int_or("height");
  if (width == -1 && height == -1) return Result::InvalidManifest;
  if (width != -1 && height != -1 && width != height) return Result::InvalidManifest;
  const long long d = width == -1 ? height : width;
  if (d < min_dim || d > Context::kMaxDimension) return Result::InvalidManifest;
  dim = (int)d;
  const std::vector<json::Json> none;
  const std::vector<json::Json>& items = tex->kind == json::Json::Array ? tex->arr : none;
  if (tex->kind != json::Json::Array && tex->kind != json::Json::Null) return Result::InvalidManifest;
  // the directory of the manifest: relative image paths are tried as given
  // (the reference opens them relative to the working directory) and then
  // relative to it
  const size_t slash = filepath.find_last_of('/');
  const std::string base = slash == std::string::npos ? "" : filepath.substr(0, slash + 1);
  for (const json::Json& t : items)
  {
    const json::Json* sems = t.find("semantics");
    const json::Json* file = t.find("fileName");
    if (!sems || sems->kind != json::Json::Object) return Result::InvalidManifest;
    if (!file || file->kind != json::Json::String) return Result::InvalidManifest;
    std::map<std::string, const json::Json*> sorted;
    for (const auto& kv : sems->obj) sorted[kv.first] = &kv.second;
    for (const auto& kv : sorted)
    {
      if (kv.second->kind != json::Json::String) return Result::InvalidManifest;
      ManifestEntry e;
      e.semantic = SemanticFromName(kv.first);
      e.channels = kv.second->str;
      if (!ValidSemanticChannels(e.semantic, e.channels)) return Result::InvalidManifest;
      if (e.semantic != Semantic::None) e.channels.resize(GetChannelCountForSemantic(e.semantic));
      e.path = file->str;
      if (!e.path.empty() && e.path[0] != '/' && !base.empty())
      {
        std::ifstream probe(e.path, std::ios::binary);
        if (!probe) e.path = base + e.path;
      }
      const json::Json* nm = t.find("name");
      e.name = nm && nm->kind == json::Json::String ? nm->str : "";
      const json::Json* srgb = t.find("isSRGB");
      e.srgb = srgb && srgb->kind == json::Json::Bool && srgb->b;
      const json::Json* flip = t.find("verticalFlip");
      e.flip = flip && flip->kind == json::Json::Bool && flip->b;
      if ((int)out.size() >= kMaxSources) return Result::InvalidManifest;
      out.push_back(e);
    }
  }
  int total = 0;
  for (const ManifestEntry& e : out) total += (int)e.channels.size();
  if (total < 1 || total > Context::kMaxChannels) return Result::InvalidManifest;
  std::stable_sort(out.begin(), out.end(),
                   [](const ManifestEntry& a, const ManifestEntry& b) { return (int)a.semantic < (int)b.semantic; });
  if (dim & (dim - 1)) return Result::InvalidDimension;
  return Result::Success;
}

// Decodes every distinct image of the entries once; checks the side.
Result DecodeAll(const std::vector<ManifestEntry>& entries, int dim, std::map<std::string, std::vector<uint8_t>>& img)
{
  for (const ManifestEntry& e : entries)
  {
    if (img.count(e.path)) continue;
    std::vector<uint8_t> rgba;
    int w = 0, h = 0;
    if (!DecodeImage(e.path, rgba, w, h)) return Result::FileNotFound;
    if (w != h || w != dim) return Result::InvalidManifest;
    img[e.path] = std::move(rgba);
  }
  return Result::Success;
}
} // namespace

Result Context::LoadManifest(const std::string& filepath)
{
  UnloadManifest();
  if (!impl_->initialized) return Result::InvalidState;
  int dim = 0;
  std::vector<ManifestEntry> entries;
  Result r = ParseManifest(filepath, impl_->min_dimension, dim, entries);
  if (r != Result::Success) return r;
  std::map<std::string, std::vector<uint8_t>> images;
  r = DecodeAll(entries, dim, images);
  if (r != Result::Success) return r;
  std::vector<TextureInput> in(entries.size());
  for (size_t i = 0; i < entries.size(); ++i)
    in[i] = TextureInput{images[entries[i].path].data(), entries[i].semantic, entries[i].channels.c_str(),
                         entries[i].name.c_str()};
  r = LoadTextures(in.data(), (int32_t)in.size(), dim);
  if (r != Result::Success) return r;
  // entries are already in the stable semantic order LoadTextures applies
  for (size_t i = 0; i < entries.size(); ++i)
  {
    TextureSource& s = impl_->manifest.sources_[i];
    s.path_ = entries[i].path;
    s.is_srgb_ = entries[i].srgb;
    s.vertical_flip_ = entries[i].flip;
  }
  return Result::Success;
}

extern "C" {
int ntc_image_decode_rgba8(const char* path, int32_t* width, int32_t* height, uint8_t* rgba, size_t capacity)
{
  if (!path || !width || !height) return (int)Result::InvalidState;
  std::vector<uint8_t> img;
  int w = 0, h = 0;
  if (!DecodeImage(path, img, w, h)) return (int)Result::FileNotFound;
  *width = w;
  *height = h;
  if (rgba)
  {
    if (capacity < img.size()) return (int)Result::InvalidState;
    memcpy(rgba, img.data(), img.size());
  }
  return 0;
}

int ntc_manifest_parse(const char* path, int32_t min_dimension, int32_t* dim, int32_t* count,
                       ntc_manifest_source* sources)
{
  if (!path || !dim || !count) return (int)Result::InvalidState;
  int d = 0;
  std::vector<ManifestEntry> entries;
  Result r = ParseManifest(path, min_dimension > 0 ? min_dimension : Context::kMinDimension, d, entries);
  *dim = d;
  *count = (int32_t)entries.size();
  if (sources)
    for (size_t i = 0; i < entries.size(); ++i)
    {
      ntc_manifest_source& s = sources[i];
      memset(&s, 0, sizeof(s));
      snprintf(s.name, sizeof(s.name), "%s", entries[i].name.c_str());
      snprintf(s.path, sizeof(s.path), "%s", entries[i].path.c_str());
      snprintf(s.channels, sizeof(s.channels), "%s", entries[i].channels.c_str());
      s.semantic = (int32_t)entries[i].semantic;
      s.is_srgb = entries[i].srgb;
      s.vertical_flip = entries[i].flip;
    }
  if (r != Result::Success) return (int)r;
  std::map<std::string, std::vector<uint8_t>> images;
  return (int)DecodeAll(entries, d, images);
}
} // extern "C"
