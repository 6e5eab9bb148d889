// JPEG and BMP decoding to 8-bit RGBA for the texture loader
// (libopenntc.cpp:1302 reads textures with stb_image's stbi_load, 4 channels).
//
// JPEG: ITU-T T.81 Huffman-coded frames, sequential (SOF0, SOF1) and
// progressive (SOF2), 8-bit samples, one (grey) or three components with any
// sampling factors that divide the largest, restart intervals. Coefficients
// of all scans are collected per component, then dequantised and inverse
// transformed (separable float IDCT); chroma planes sampled at half
// resolution are brought to full size with the triangle filter stb_image
// applies (weights 3/4, 1/4 along each halved axis, libjpeg's "fancy"
// upsampling), other ratios by replication; YCbCr (or RGB when the
// components are tagged 'R', 'G', 'B' or an Adobe marker says so) to RGB in
// float with rounding. Arithmetic coding, 12-bit samples and four-component
// (CMYK) images are not read (decoding fails, as a failed load).
//
// BMP: BITMAPINFOHEADER and later headers and the OS/2 header, 1/4/8-bit
// palettes, 16/24/32-bit pixels with default or bit-field masks (uncompressed
// or BI_BITFIELDS; run-length compression is not read), bottom-up or top-down
// rows; channels narrower than 8 bits widened by bit replication; a 32-bit
// image whose alpha is zero everywhere reads as opaque.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "images.h"

namespace openntc
{
namespace img
{
namespace
{
// natural (row-major) position of the k-th coefficient in zig-zag order (T.81 Figure A.6)
const uint8_t kNatural[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                              12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                              35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                              58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 Annex C / F.2.2.3), with a 9-bit lookup.
struct Huffman
{
  uint8_t vals[256] = {};
  int32_t maxcode[18] = {};
  int32_t valptr[17] = {};
  int32_t mincode[17] = {};
  uint16_t fast[512] = {};  // (length << 8) | value for codes of <= 9 bits; 0: longer
  bool defined = false;

  bool Build(const uint8_t counts[16], const uint8_t* symbols, int n)
  {
    if (n > 256) return false;
    std::memcpy(vals, symbols, (size_t)n);
    std::memset(fast, 0, sizeof(fast));
    int code = 0, k = 0;
    for (int l = 1; l <= 16; ++l)
    {
      valptr[l] = k;
      mincode[l] = code;
      for (int i = 0; i < counts[l - 1]; ++i, ++k, ++code)
        if (l <= 9)
        {
          const int shift = 9 - l;
          for (int j = 0; j < (1 << shift); ++j) fast[(code << shift) | j] = (uint16_t)((l << 8) | vals[k]);
        }
      maxcode[l] = counts[l - 1] ? code - 1 : -1;
      if (code > (1 << l)) return false;
      code <<= 1;
    }
    maxcode[17] = 0x7fffffff;
    defined = true;
    return true;
  }
};

// Entropy-coded data reader: byte stuffing (FF 00) removed; at a marker it
// stops consuming and feeds zero bits.
struct Bits
{
  const uint8_t* p = nullptr;
  const uint8_t* end = nullptr;
  uint64_t buf = 0;
  int n = 0;
  bool at_marker = false;

  void Fill()
  {
    while (n <= 56)
    {
      uint8_t b = 0;
      if (!at_marker && p < end)
      {
        b = *p;
        if (b == 0xFF)
        {
          if (p + 1 < end && p[1] == 0x00)
            p += 2;
          else
          {
            at_marker = true;
            b = 0;
          }
        }
        else
          ++p;
      }
      buf |= (uint64_t)b << (56 - n);
      n += 8;
    }
  }
  uint32_t Peek(int k)
  {
    if (n < k) Fill();
    return (uint32_t)(buf >> (64 - k));
  }
  void Skip(int k)
  {
    buf <<= k;
    n -= k;
  }
  uint32_t Get(int k)
  {
    if (k == 0) return 0;
    const uint32_t v = Peek(k);
    Skip(k);
    return v;
  }
  int Decode(const Huffman& h)
  {
    const uint32_t top = Peek(16);
    const uint16_t f = h.fast[top >> 7];
    if (f)
    {
      Skip(f >> 8);
      return f & 0xFF;
    }
    for (int l = 10; l <= 16; ++l)
    {
      const int32_t code = (int32_t)(top >> (16 - l));
      if (code <= h.maxcode[l])
      {
        Skip(l);
        const int idx = h.valptr[l] + code - h.mincode[l];
        return idx >= 0 && idx < 256 ? h.vals[idx] : -1;
      }
    }
    return -1;
  }
  // T.81 F.2.2.1 EXTEND of an s-bit magnitude category
  int Receive(int s)
  {
    if (s == 0) return 0;
    const int v = (int)Get(s);
    return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
  }
  // restart: drop the partial byte, step over the RSTn marker
  void Restart()
  {
    buf = 0;
    n = 0;
    at_marker = false;
    while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) ++p;
    if (p + 1 < end) p += 2;
  }
};

struct Component
{
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;    // Huffman tables of the current scan
  int bw = 0, bh = 0;    // blocks per row / column (padded to whole MCUs)
  int cw = 0, ch = 0;    // samples per row / column
  std::vector<int16_t> coef;  // [bh][bw][64] natural order
  int dc_pred = 0;
};

struct Jpeg
{
  uint16_t q[4][64] = {};  // natural order
  Huffman dc[4], ac[4];
  std::vector<Component> comp;
  int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
  bool progressive = false, frame = false, jfif = false;
  int adobe = -1;  // colour transform of an Adobe APP14 marker
  int restart = 0;
  int eobrun = 0;

  int16_t* Block(Component& c, int bx, int by) { return &c.coef[((size_t)by * c.bw + bx) * 64]; }

  bool DecodeBlock(Bits& b, Component& c, int16_t* blk, int ss, int se, int ah, int al)
  {
    if (!progressive)
    {
      const int t = b.Decode(dc[c.td]);
      if (t < 0 || t > 16) return false;
      c.dc_pred += b.Receive(t);
      blk[0] = (int16_t)c.dc_pred;
      for (int k = 1; k < 64;)
      {
        const int rs = b.Decode(ac[c.ta]);
        if (rs < 0) return false;
        const int r = rs >> 4, s = rs & 15;
        if (s == 0)
        {
          if (r != 15) break;
          k += 16;
          continue;
        }
        k += r;
        if (k > 63) return false;
        blk[kNatural[k]] = (int16_t)b.Receive(s);
        ++k;
      }
      return true;
    }
    if (ss == 0)  // DC scans
    {
      if (ah == 0)
      {
        const int t = b.Decode(dc[c.td]);
        if (t < 0 || t > 16) return false;
        c.dc_pred += b.Receive(t);
        blk[0] = (int16_t)(c.dc_pred * (1 << al));
      }
      else if (b.Get(1))
        blk[0] = (int16_t)(blk[0] | (1 << al));
      return true;
    }
    if (ah == 0)  // AC first pass (T.81 G.1.2.2)
    {
      if (eobrun > 0)
      {
        --eobrun;
        return true;
      }
      for (int k = ss; k <= se; ++k)
      {
        const int rs = b.Decode(ac[c.ta]);
        if (rs < 0) return false;
        const int r = rs >> 4, s = rs & 15;
        if (s == 0)
        {
          if (r < 15)
          {
            eobrun = (1 << r) - 1;
            if (r) eobrun += (int)b.Get(r);
            break;
          }
          k += 15;
          continue;
        }
        k += r;
        if (k > 63) return false;
        blk[kNatural[k]] = (int16_t)(b.Receive(s) * (1 << al));
      }
      return true;
    }
    // AC refinement (T.81 G.1.2.3): a correction bit for every coefficient
    // already nonzero on the way, new coefficients of magnitude 1 << al
    const int p1 = 1 << al, m1 = -(1 << al);
    int k = ss;
    if (eobrun == 0)
    {
      for (; k <= se; ++k)
      {
        const int rs = b.Decode(ac[c.ta]);
        if (rs < 0) return false;
        int r = rs >> 4, s = rs & 15, val = 0;
        if (s)
          val = b.Get(1) ? p1 : m1;
        else if (r != 15)
        {
          eobrun = 1 << r;
          if (r) eobrun += (int)b.Get(r);
          break;
        }
        while (k <= se)
        {
          int16_t& cf = blk[kNatural[k]];
          if (cf != 0)
          {
            if (b.Get(1) && (cf & p1) == 0) cf = (int16_t)(cf >= 0 ? cf + p1 : cf + m1);
          }
          else if (--r < 0)
            break;
          ++k;
        }
        if (s && k <= se) blk[kNatural[k]] = (int16_t)val;
      }
    }
    if (eobrun > 0)
    {
      for (; k <= se; ++k)
      {
        int16_t& cf = blk[kNatural[k]];
        if (cf != 0 && b.Get(1) && (cf & p1) == 0) cf = (int16_t)(cf >= 0 ? cf + p1 : cf + m1);
      }
      --eobrun;
    }
    return true;
  }

  bool Scan(const uint8_t*& p, const uint8_t* end, const std::vector<int>& ci, int ss, int se, int ah, int al)
  {
    Bits b;
    b.p = p;
    b.end = end;
    eobrun = 0;
    for (Component& c : comp) c.dc_pred = 0;
    int todo = restart;
    // after a unit that is not the scan's last: count towards the restart interval
    auto next_unit = [&]() {
      if (restart && --todo == 0)
      {
        b.Restart();
        todo = restart;
        eobrun = 0;
        for (Component& c : comp) c.dc_pred = 0;
      }
    };
    if (ci.size() == 1)
    {
      // one component: its blocks in raster order over its own size
      Component& c = comp[ci[0]];
      const int w = (c.cw + 7) / 8, h = (c.ch + 7) / 8;
      for (int by = 0; by < h; ++by)
        for (int bx = 0; bx < w; ++bx)
        {
          if (!DecodeBlock(b, c, Block(c, bx, by), ss, se, ah, al)) return false;
          if (by + 1 < h || bx + 1 < w) next_unit();
        }
    }
    else
    {
      for (int my = 0; my < mcuy; ++my)
        for (int mx = 0; mx < mcux; ++mx)
        {
          for (int i : ci)
          {
            Component& c = comp[i];
            for (int y = 0; y < c.v; ++y)
              for (int x = 0; x < c.h; ++x)
                if (!DecodeBlock(b, c, Block(c, mx * c.h + x, my * c.v + y), ss, se, ah, al)) return false;
          }
          if (my + 1 < mcuy || mx + 1 < mcux) next_unit();
        }
    }
    // continue at the next marker
    p = b.p;
    while (p + 1 < end && !(p[0] == 0xFF && p[1] != 0x00 && !(p[1] >= 0xD0 && p[1] <= 0xD7))) ++p;
    return true;
  }
};

uint16_t Be16(const uint8_t* p) { return (uint16_t)((p[0] << 8) | p[1]); }

// 8x8 inverse DCT (T.81 A.3.3) in float, level shift, rounding and clamping.
struct IdctTable
{
  float c[8][8];  // c[x][u] = C(u) / 2 cos((2x + 1) u pi / 16), C(0) = 1 / sqrt 2
  IdctTable()
  {
    const double pi = std::acos(-1.0);
    for (int x = 0; x < 8; ++x)
      for (int u = 0; u < 8; ++u)
        c[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * pi / 16.0));
  }
};

void Idct(const int16_t* in, const uint16_t* q, uint8_t* out, int stride)
{
  static const IdctTable T;
  const auto& c = T.c;
  float f[64], t[64];
  for (int i = 0; i < 64; ++i) f[i] = (float)in[i] * (float)q[i];
  for (int y = 0; y < 8; ++y)
    for (int x = 0; x < 8; ++x)
    {
      float s = 0.0f;
      for (int u = 0; u < 8; ++u) s += c[x][u] * f[y * 8 + u];
      t[y * 8 + x] = s;
    }
  for (int x = 0; x < 8; ++x)
    for (int y = 0; y < 8; ++y)
    {
      float s = 0.0f;
      for (int v = 0; v < 8; ++v) s += c[y][v] * t[v * 8 + x];
      const float r = std::floor(s + 128.5f);
      out[y * stride + x] = (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
    }
}

// One full-resolution row of a chroma plane subsampled by hs x vs (the
// triangle filter for factors of 2, replication otherwise).
void UpsampleRow(const uint8_t* plane, int stride, int cw, int ch, int hs, int vs, int y, int width, uint8_t* out)
{
  const int cy = y / vs;
  const uint8_t* near_row = plane + (size_t)cy * stride;
  const uint8_t* far_row = near_row;
  if (vs == 2)
  {
    // even rows lean on the row above, odd rows on the row below (clamped)
    const int fy = (y & 1) ? cy + 1 : cy - 1;
    far_row = plane + (size_t)(fy < 0 ? 0 : (fy >= ch ? ch - 1 : fy)) * stride;
  }
  std::vector<int> t(cw);
  if (vs == 2)
    for (int i = 0; i < cw; ++i) t[i] = 3 * near_row[i] + far_row[i];
  if (hs == 1 && vs == 1)
    std::memcpy(out, near_row, (size_t)width);
  else if (hs == 1 && vs == 2)
    for (int i = 0; i < width; ++i) out[i] = (uint8_t)((t[i] + 2) >> 2);
  else if (hs == 2 && vs == 1)
  {
    std::vector<uint8_t> o(2 * cw);
    if (cw == 1)
      o[0] = o[1] = near_row[0];
    else
    {
      o[0] = near_row[0];
      o[1] = (uint8_t)((near_row[0] * 3 + near_row[1] + 2) >> 2);
      for (int i = 1; i < cw - 1; ++i)
      {
        const int n = 3 * near_row[i] + 2;
        o[2 * i] = (uint8_t)((n + near_row[i - 1]) >> 2);
        o[2 * i + 1] = (uint8_t)((n + near_row[i + 1]) >> 2);
      }
      o[2 * cw - 2] = (uint8_t)((near_row[cw - 2] * 3 + near_row[cw - 1] + 2) >> 2);
      o[2 * cw - 1] = near_row[cw - 1];
    }
    std::memcpy(out, o.data(), (size_t)width);
  }
  else if (hs == 2 && vs == 2)
  {
    std::vector<uint8_t> o(2 * cw);
    if (cw == 1)
      o[0] = o[1] = (uint8_t)((t[0] + 2) >> 2);
    else
    {
      o[0] = (uint8_t)((t[0] + 2) >> 2);
      for (int i = 1; i < cw; ++i)
      {
        o[2 * i - 1] = (uint8_t)((3 * t[i - 1] + t[i] + 8) >> 4);
        o[2 * i] = (uint8_t)((3 * t[i] + t[i - 1] + 8) >> 4);
      }
      o[2 * cw - 1] = (uint8_t)((t[cw - 1] + 2) >> 2);
    }
    std::memcpy(out, o.data(), (size_t)width);
  }
  else
    for (int i = 0; i < width; ++i) out[i] = near_row[i / hs];
}
} // namespace

bool DecodeJpeg(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h)
{
  const uint8_t* p = f.data();
  const uint8_t* end = p + f.size();
  if (f.size() < 4 || p[0] != 0xFF || p[1] != 0xD8) return false;
  p += 2;
  Jpeg J;
  for (;;)
  {
    while (p < end && *p != 0xFF) ++p;  // stray bytes before a marker
    while (p < end && *p == 0xFF) ++p;  // fill bytes
    if (p >= end) break;
    const uint8_t m = *p++;
    if (m == 0xD9) break;                          // EOI
    if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;  // TEM, stray RSTn
    if (p + 2 > end) return false;
    const int len = Be16(p);
    if (len < 2 || p + len > end) return false;
    const uint8_t* s = p + 2;
    const uint8_t* se = p + len;
    if (m == 0xDB)  // DQT
    {
      while (s < se)
      {
        const int pq = s[0] >> 4, tq = s[0] & 15;
        ++s;
        if (tq > 3 || s + (pq ? 128 : 64) > se) return false;
        for (int k = 0; k < 64; ++k) J.q[tq][kNatural[k]] = pq ? Be16(s + 2 * k) : s[k];
        s += pq ? 128 : 64;
      }
    }
    else if (m == 0xC4)  // DHT
    {
      while (s < se)
      {
        if (s + 17 > se) return false;
        const int tc = s[0] >> 4, th = s[0] & 15;
        if (tc > 1 || th > 3) return false;
        int n = 0;
        for (int i = 0; i < 16; ++i) n += s[1 + i];
        if (s + 17 + n > se) return false;
        if (!(tc ? J.ac[th] : J.dc[th]).Build(s + 1, s + 17, n)) return false;
        s += 17 + n;
      }
    }
    else if (m == 0xC0 || m == 0xC1 || m == 0xC2)  // SOF: Huffman, sequential or progressive
    {
      if (J.frame || len < 8 || s[0] != 8) return false;
      J.progressive = m == 0xC2;
      J.height = Be16(s + 1);
      J.width = Be16(s + 3);
      const int nc = s[5];
      if ((nc != 1 && nc != 3) || len < 8 + 3 * nc || !SaneImageSize(J.width, J.height)) return false;
      J.comp.resize(nc);
      for (int i = 0; i < nc; ++i)
      {
        Component& c = J.comp[i];
        c.id = s[6 + 3 * i];
        c.h = s[7 + 3 * i] >> 4;
        c.v = s[7 + 3 * i] & 15;
        c.tq = s[8 + 3 * i];
        if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return false;
        J.hmax = std::max(J.hmax, c.h);
        J.vmax = std::max(J.vmax, c.v);
      }
      J.mcux = (J.width + 8 * J.hmax - 1) / (8 * J.hmax);
      J.mcuy = (J.height + 8 * J.vmax - 1) / (8 * J.vmax);
      for (Component& c : J.comp)
      {
        if (J.hmax % c.h || J.vmax % c.v) return false;
        c.bw = J.mcux * c.h;
        c.bh = J.mcuy * c.v;
        c.cw = (J.width * c.h + J.hmax - 1) / J.hmax;
        c.ch = (J.height * c.v + J.vmax - 1) / J.vmax;
        c.coef.assign((size_t)c.bw * c.bh * 64, 0);
      }
      J.frame = true;
    }
    else if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC)
      return false;  // lossless, hierarchical or arithmetic-coded frames
    else if (m == 0xDD)  // DRI
    {
      if (len < 4) return false;
      J.restart = Be16(s);
    }
    else if (m == 0xDA)  // SOS
    {
      if (!J.frame || len < 6) return false;
      const int ns = s[0];
      if (ns < 1 || ns > 4 || len < 6 + 2 * ns) return false;
      std::vector<int> ci;
      for (int i = 0; i < ns; ++i)
      {
        const int id = s[1 + 2 * i], tables = s[2 + 2 * i];
        int found = -1;
        for (size_t k = 0; k < J.comp.size(); ++k)
          if (J.comp[k].id == id) found = (int)k;
        if (found < 0) return false;
        J.comp[found].td = tables >> 4;
        J.comp[found].ta = tables & 15;
        if (J.comp[found].td > 3 || J.comp[found].ta > 3) return false;
        ci.push_back(found);
      }
      const int ss = s[1 + 2 * ns], sse = s[2 + 2 * ns], ah = s[3 + 2 * ns] >> 4, al = s[3 + 2 * ns] & 15;
      if (J.progressive)
      {
        if (ss > sse || sse > 63 || (ss == 0 && sse != 0) || (ss > 0 && ns != 1) || al > 13) return false;
      }
      else if (ss != 0 || sse != 63)
        return false;
      for (int i : ci)
      {
        const Component& c = J.comp[i];
        if ((ss == 0 && ah == 0 && !J.dc[c.td].defined) || (sse > 0 && !J.ac[c.ta].defined)) return false;
      }
      p = se;
      if (!J.Scan(p, end, ci, ss, sse, ah, al)) return false;
      continue;
    }
    else if (m == 0xE0 && len >= 7 && std::memcmp(s, "JFIF", 5) == 0)
      J.jfif = true;
    else if (m == 0xEE && len >= 14 && std::memcmp(s, "Adobe", 5) == 0)
      J.adobe = s[11];
    // other segments (APPn, COM, DNL) carry nothing needed here
    p = se;
  }
  if (!J.frame) return false;

  // dequantise and transform every block into the component planes
  const int W = J.width, H = J.height;
  std::vector<std::vector<uint8_t>> plane(J.comp.size());
  for (size_t i = 0; i < J.comp.size(); ++i)
  {
    Component& c = J.comp[i];
    const int stride = c.bw * 8;
    plane[i].assign((size_t)stride * c.bh * 8, 0);
    for (int by = 0; by < c.bh; ++by)
      for (int bx = 0; bx < c.bw; ++bx)
        Idct(J.Block(c, bx, by), J.q[c.tq], &plane[i][(size_t)by * 8 * stride + bx * 8], stride);
  }
  w = W;
  h = H;
  rgba.assign((size_t)W * H * 4, 255);
  if (J.comp.size() == 1)
  {
    const Component& c = J.comp[0];
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x)
      {
        const uint8_t v = plane[0][(size_t)y * c.bw * 8 + x];
        uint8_t* o = &rgba[((size_t)y * W + x) * 4];
        o[0] = o[1] = o[2] = v;
      }
    return true;
  }
  const bool rgb = (J.comp[0].id == 'R' && J.comp[1].id == 'G' && J.comp[2].id == 'B') || (J.adobe == 0 && !J.jfif);
  std::vector<uint8_t> row[3];
  for (int k = 0; k < 3; ++k) row[k].resize((size_t)W + 8);
  for (int y = 0; y < H; ++y)
  {
    for (int k = 0; k < 3; ++k)
    {
      const Component& c = J.comp[k];
      UpsampleRow(plane[k].data(), c.bw * 8, c.cw, c.ch, J.hmax / c.h, J.vmax / c.v, y, W, row[k].data());
    }
    for (int x = 0; x < W; ++x)
    {
      uint8_t* o = &rgba[((size_t)y * W + x) * 4];
      if (rgb)
      {
        o[0] = row[0][x];
        o[1] = row[1][x];
        o[2] = row[2][x];
        continue;
      }
      // JFIF YCbCr -> RGB
      const float Y = row[0][x], cb = (float)row[1][x] - 128.0f, cr = (float)row[2][x] - 128.0f;
      const float r = Y + 1.402f * cr, g = Y - 0.344136f * cb - 0.714136f * cr, b = Y + 1.772f * cb;
      auto cl = [](float v) {
        const float t = std::floor(v + 0.5f);
        return (uint8_t)(t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t));
      };
      o[0] = cl(r);
      o[1] = cl(g);
      o[2] = cl(b);
    }
  }
  return true;
}

namespace
{
uint32_t Le32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
uint16_t Le16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }

// channel under mask, widened to 8 bits by bit replication
uint8_t Channel(uint32_t v, uint32_t mask)
{
  if (!mask) return 0;
  int shift = 0, bits = 0;
  while (!((mask >> shift) & 1)) ++shift;
  while (shift + bits < 32 && ((mask >> (shift + bits)) & 1)) ++bits;
  uint32_t x = (v & mask) >> shift;
  if (bits >= 8) return (uint8_t)(x >> (bits - 8));
  uint32_t out = 0;
  int filled = 0;
  while (filled < 8)
  {
    out = (out << bits) | x;
    filled += bits;
  }
  return (uint8_t)(out >> (filled - 8));
}
} // namespace

bool DecodeBmp(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h)
{
  if (f.size() < 26 || f[0] != 'B' || f[1] != 'M') return false;
  const uint8_t* d = f.data();
  const uint32_t offset = Le32(d + 10), hsz = Le32(d + 14);
  if (hsz != 12 && hsz != 40 && hsz != 52 && hsz != 56 && hsz != 108 && hsz != 124) return false;
  if (14 + (size_t)hsz > f.size()) return false;
  int32_t width, height;
  int bpp, compression = 0;
  if (hsz == 12)
  {
    width = Le16(d + 18);
    height = (int16_t)Le16(d + 20);
    bpp = Le16(d + 24);
  }
  else
  {
    width = (int32_t)Le32(d + 18);
    height = (int32_t)Le32(d + 22);
    bpp = Le16(d + 28);
    compression = (int)Le32(d + 30);
  }
  if (compression != 0 && compression != 3) return false;  // run-length and other codings
  if (height == INT32_MIN) return false;
  const bool top_down = height < 0;
  if (top_down) height = -height;
  if (!SaneImageSize(width, height)) return false;
  uint32_t mr = 0, mg = 0, mb = 0, ma = 0;
  if (bpp == 16)
  {
    mr = 31u << 10;
    mg = 31u << 5;
    mb = 31u;
  }
  else if (bpp == 32)
  {
    mr = 0xffu << 16;
    mg = 0xffu << 8;
    mb = 0xffu;
    ma = 0xffu << 24;
  }
  if (compression == 3)
  {
    if (bpp != 16 && bpp != 32) return false;
    if (hsz == 40)
    {
      if (54 + 12 > f.size()) return false;
      mr = Le32(d + 54);
      mg = Le32(d + 58);
      mb = Le32(d + 62);
      ma = 0;
    }
    else
    {
      mr = Le32(d + 54);
      mg = Le32(d + 58);
      mb = Le32(d + 62);
      ma = Le32(d + 66);
    }
    if (!mr || !mg || !mb) return false;
  }
  else if (bpp == 32 && hsz >= 56)
  {
    // later headers carry the alpha mask even for uncompressed data
    const uint32_t a = Le32(d + 66);
    if (a) ma = a;
  }
  std::vector<uint8_t> pal;  // BGR(x) entries
  if (bpp <= 8)
  {
    if (bpp != 1 && bpp != 4 && bpp != 8) return false;
    const int entry = hsz == 12 ? 3 : 4;
    const size_t start = 14 + hsz;
    if (offset < start) return false;
    const size_t n = (offset - start) / entry;
    if (n == 0 || start + n * entry > f.size()) return false;
    pal.assign(256 * 4, 0);
    for (size_t i = 0; i < n && i < 256; ++i)
      for (int k = 0; k < 3; ++k) pal[i * 4 + k] = d[start + i * entry + k];
  }
  else if (bpp != 16 && bpp != 24 && bpp != 32)
    return false;
  const size_t stride = (((size_t)width * bpp + 31) / 32) * 4;
  if ((size_t)offset + stride * height > f.size()) return false;
  w = width;
  h = height;
  rgba.assign((size_t)width * height * 4, 255);
  bool any_alpha = false;
  for (int r = 0; r < height; ++r)
  {
    const uint8_t* src = d + offset + stride * r;
    const int y = top_down ? r : height - 1 - r;
    uint8_t* out = &rgba[(size_t)y * width * 4];
    for (int x = 0; x < width; ++x)
    {
      uint8_t* o = out + 4 * x;
      if (bpp <= 8)
      {
        int idx;
        if (bpp == 8)
          idx = src[x];
        else if (bpp == 4)
          idx = (src[x >> 1] >> ((x & 1) ? 0 : 4)) & 15;
        else
          idx = (src[x >> 3] >> (7 - (x & 7))) & 1;
        o[0] = pal[idx * 4 + 2];
        o[1] = pal[idx * 4 + 1];
        o[2] = pal[idx * 4 + 0];
      }
      else if (bpp == 24)
      {
        o[0] = src[3 * x + 2];
        o[1] = src[3 * x + 1];
        o[2] = src[3 * x + 0];
      }
      else
      {
        const uint32_t v = bpp == 16 ? Le16(src + 2 * x) : Le32(src + 4 * x);
        o[0] = Channel(v, mr);
        o[1] = Channel(v, mg);
        o[2] = Channel(v, mb);
        if (ma)
        {
          o[3] = Channel(v, ma);
          any_alpha = any_alpha || o[3] != 0;
        }
      }
    }
  }
  if (ma && !any_alpha)
    for (size_t i = 3; i < rgba.size(); i += 4) rgba[i] = 255;
  return true;
}
} // namespace img
} // namespace openntc
