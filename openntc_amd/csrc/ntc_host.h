// Host-side helpers shared by the C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ntc_hip.h"

#define NTC_RETURN_IF(cond, code) \
  do                              \
  {                               \
    if (cond) return (code);      \
  } while (0)

#define NTC_HIP_TRY(expr)                       \
  do                                            \
  {                                             \
    hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return (int)e_;       \
  } while (0)

// Status of the most recent launch on this thread.
#define NTC_LAUNCH_STATUS() ((int)hipGetLastError())

namespace ntc
{
inline unsigned blocks_for(int64_t n, int threads)
{
  return (unsigned)((n + threads - 1) / threads);
}

// Feature-vector slot bookkeeping (matches the gather's write order).
inline int g1_base(const ntc_feature_desc& fd)
{
  return fd.g0_mode == NTC_G0_CORNERS ? 4 * fd.g0_channels : fd.g0_channels;
}

inline int used_slots(const ntc_feature_desc& fd)
{
  return g1_base(fd) + fd.g1_channels + 4 * fd.pe_count + (fd.lod_slot ? 1 : 0);
}

inline bool valid_feature_desc(const ntc_feature_desc* fd)
{
  if (!fd) return false;
  if (fd->g0_dim < 1 || fd->g1_dim < 1) return false;
  if (fd->g0_channels < 1 || fd->g0_channels > 32) return false;
  if (fd->g1_channels < 1 || fd->g1_channels > 32) return false;
  if (fd->g0_mode != NTC_G0_CORNERS && fd->g0_mode != NTC_G0_BILINEAR) return false;
  if (fd->pe_count < 0 || fd->pe_count > 4) return false;
  if (fd->pe_count > 0 && (fd->pe_first_period >> (fd->pe_count - 1)) < 1) return false;
  if (used_slots(*fd) > fd->feature_dim || fd->feature_dim > 128) return false;
  return true;
}

inline bool valid_tiles(const ntc_tiles* t)
{
  if (!t || !t->origins) return false;
  if (t->n_tiles < 1 || t->n_tiles > 64) return false;
  if (t->tile_dim < 1 || t->mip_dim < 1 || t->tile_dim > t->mip_dim) return false;
  return true;
}

inline int64_t batch_of(const ntc_tiles& t)
{
  return (int64_t)t.n_tiles * t.tile_dim * t.tile_dim;
}
} // namespace ntc
