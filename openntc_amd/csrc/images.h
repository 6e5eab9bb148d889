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
bool DecodeJpeg(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h);
bool DecodeBmp(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgba, int& w, int& h);
} // namespace img
} // namespace openntc
