// JPEG and BMP decoders of the texture loader (images.cpp): 8-bit RGBA,
// rows top to bottom; false when the data is not such an image or uses a
// form the decoders do not read.
#pragma once

#include <cstdint>
#include <vector>

namespace openntc
{
namespace img
{
// Largest image the decoders accept: 2^14 per side and 2^28 pixels (1 GiB of
// RGBA8). Headers are checked against it before anything is sized from
// them, so a corrupt header fails the decode instead of the allocation; the
// texture sets themselves are at most 8192^2.
constexpr int64_t kMaxImageSide = (int64_t)1 << 14;
constexpr int64_t kMaxImagePixels = (int64_t)1 << 28;
inline bool SaneImageSize(int64_t w, int64_t h)
{
  return w >= 1 && h >= 1 && w <= kMaxImageSide && h <= kMaxImageSide && w * h <= kMaxImagePixels;
}

bool DecodeJpeg(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h);
bool DecodeBmp(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h);
} // namespace img
} // namespace openntc
