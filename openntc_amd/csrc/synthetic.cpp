// Seeded synthetic 9-channel PBR texture set (in-memory input for training
// and benchmarks; no image files involved).
//
//   albedo RGB : three octaves of value noise per channel in [0,1], passed
//                through the sRGB decode curve, stored as uint8
//   normal RGB : unit normals of a 3-octave value-noise height field
//                (central differences), mapped from [-1,1] to [0,1]
//   orm        : R = ambient occlusion, G = roughness, B = metallic, each an
//                independent value-noise field
// Value noise: hashed lattice values (integer hash "lowbias32", C. Wellons)
// interpolated with a smoothstep-weighted bilinear blend, lattice wrapping at
// the texture edge so the set tiles.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "ntc_hip.h"

namespace
{
inline uint32_t lowbias32(uint32_t x)
{
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

inline float lattice(uint32_t seed, uint32_t field, int ix, int iy)
{
  const uint32_t h = lowbias32(seed ^ lowbias32(field * 0x9E3779B9u ^ lowbias32((uint32_t)ix * 0x85EBCA6Bu ^
                                                                                   (uint32_t)iy * 0xC2B2AE35u)));
  return (float)(h >> 8) * 0x1p-24f;
}

inline float smooth(float t) { return t * t * (3.0f - 2.0f * t); }

// One octave: lattice period `cells` across a texture of side dim.
inline float value_noise(uint32_t seed, uint32_t field, int dim, int cells, int x, int y)
{
  const float s = (float)cells / (float)dim;
  const float gx = ((float)x + 0.5f) * s, gy = ((float)y + 0.5f) * s;
  const int x0 = (int)std::floor(gx), y0 = (int)std::floor(gy);
  const float fx = smooth(gx - (float)x0), fy = smooth(gy - (float)y0);
  auto L = [&](int ix, int iy) { return lattice(seed, field, ((ix % cells) + cells) % cells, ((iy % cells) + cells) % cells); };
  const float a = L(x0, y0), b = L(x0 + 1, y0), c = L(x0, y0 + 1), d = L(x0 + 1, y0 + 1);
  return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy;
}

// Three octaves, cells = base, 2 base, 4 base with amplitudes 1/2, 1/4, 1/8
// renormalised to [0,1].
inline float fractal(uint32_t seed, uint32_t field, int dim, int base, int x, int y)
{
  float v = 0.0f, amp = 0.5f, norm = 0.0f;
  int cells = base;
  for (int o = 0; o < 3; ++o)
  {
    v += amp * value_noise(seed, field * 8u + (uint32_t)o, dim, std::min(cells, dim), x, y);
    norm += amp;
    amp *= 0.5f;
    cells *= 2;
  }
  return v / norm;
}

inline uint8_t to_u8(float v)
{
  v = std::min(1.0f, std::max(0.0f, v));
  return (uint8_t)std::lround(v * 255.0f);
}

inline float srgb_decode(float c)
{
  return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
}

template <typename Fn>
void parallel_rows(int dim, Fn&& fn)
{
  const unsigned n = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < n; ++t)
    th.emplace_back([&, t] {
      for (int y = (int)t; y < dim; y += (int)n) fn(y);
    });
  for (auto& t : th) t.join();
}
} // namespace

extern "C" int ntc_synthetic_pbr(int32_t dim, uint32_t seed, uint8_t* albedo, uint8_t* normal, uint8_t* orm)
{
  if (dim < 4 || (dim & (dim - 1)) != 0 || !albedo || !normal || !orm) return NTC_E_INVALID;
  const int base = std::max(2, dim / 64);
  // height field first (normals need neighbours)
  std::vector<float> h((size_t)dim * dim);
  parallel_rows(dim, [&](int y) {
    for (int x = 0; x < dim; ++x) h[(size_t)y * dim + x] = fractal(seed, 10u, dim, base, x, y);
  });
  const float strength = (float)dim / 16.0f;
  parallel_rows(dim, [&](int y) {
    for (int x = 0; x < dim; ++x)
    {
      const size_t i = (size_t)y * dim + x;
      uint8_t* a = albedo + 4 * i;
      for (int c = 0; c < 3; ++c) a[c] = to_u8(srgb_decode(fractal(seed, 1u + (uint32_t)c, dim, base, x, y)));
      a[3] = 255;
      const float hl = h[(size_t)y * dim + (x + dim - 1) % dim], hr = h[(size_t)y * dim + (x + 1) % dim];
      const float hd = h[(size_t)((y + dim - 1) % dim) * dim + x], hu = h[(size_t)((y + 1) % dim) * dim + x];
      float nx = -(hr - hl) * 0.5f * strength, ny = -(hu - hd) * 0.5f * strength, nz = 1.0f;
      const float inv = 1.0f / std::sqrt(nx * nx + ny * ny + nz * nz);
      nx *= inv;
      ny *= inv;
      nz *= inv;
      uint8_t* n = normal + 4 * i;
      n[0] = to_u8(nx * 0.5f + 0.5f);
      n[1] = to_u8(ny * 0.5f + 0.5f);
      n[2] = to_u8(nz * 0.5f + 0.5f);
      n[3] = 255;
      uint8_t* o = orm + 4 * i;
      o[0] = to_u8(fractal(seed, 20u, dim, base, x, y));  // AO
      o[1] = to_u8(fractal(seed, 21u, dim, base, x, y));  // roughness
      o[2] = to_u8(fractal(seed, 22u, dim, base, x, y));  // metallic
      o[3] = 255;
    }
  });
  return NTC_OK;
}
