// Numeric formats of the evaluation decode, the same bits on the host (table
// building) and on the device: float bit access, BF16 round to nearest even,
// and the OCP FP8 E4M3 quantise-dequantise with saturation.
#pragma once

#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>

namespace ntc
{
__host__ __device__ inline uint32_t f_bits(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

__host__ __device__ inline float f_from(uint32_t u)
{
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__host__ __device__ inline float bf16_round(float x)
{
  uint32_t u = f_bits(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return f_from(u & 0xFFFF0000u);
}

// OCP FP8 E4M3 (bias 7, largest finite 448, 0x7F NaN, no infinities) quantise-dequantise of
// v * r: round to nearest even, saturating (NaN saturates to 448 with its sign). Normal range:
// the float's mantissa rounded to 3 bits on its bits (a carry moves into the exponent; the clamp
// to 448 keeps the result finite); below 2^-6: multiples of 2^-9. Formed on the float directly,
// without an E4M3 code in between: equal, on every float bit pattern, to encoding the value to
// its E4M3 byte and decoding it again.
__host__ __device__ inline float fp8_q(float v, float r)
{
  const uint32_t u = f_bits(v * r);
  float a = f_from(u & 0x7FFFFFFFu);
  a = a < 448.0f ? a : 448.0f;
  uint32_t n = f_bits(a);
  n = (n + 0x7FFFFu + ((n >> 20) & 1u)) & 0xFFF00000u;
  const float q = a < 0x1p-6f ? rintf(a * 512.0f) * 0x1p-9f : f_from(n);
  return f_from(f_bits(q) | (u & 0x80000000u));
}


// E4M3 byte (OCP: bias 7, 0x7E = 448) of a float that fp8_q returned (an
// E4M3 value, so the encoding is exact): sign, then 2^-9 multiples below 2^-6,
// else exponent field and the 3 mantissa bits.
__host__ __device__ inline uint8_t e4m3_byte(float q)
{
  const uint32_t u = f_bits(q);
  const uint32_t s = (u >> 24) & 0x80u;
  const float a = f_from(u & 0x7FFFFFFFu);
  if (a < 0x1p-6f) return (uint8_t)(s | (uint32_t)(a * 512.0f));
  const uint32_t e = ((u >> 23) & 0xFFu) - 127u + 7u;
  return (uint8_t)(s | (e << 3) | ((u >> 20) & 7u));
}
} // namespace ntc
