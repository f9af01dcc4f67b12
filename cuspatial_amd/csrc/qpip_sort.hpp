// Device helpers of the LSD radix sort shared with the fused keygen kernel.
#pragma once

#include "qpip_internal.hpp"

namespace qpip {

// Lanes of the wave whose digit equals this lane's digit (restricted to `active`).
// One ballot per digit bit: peers = AND over bits of (ballot(bit) or its complement).
template <int NB>
__device__ __forceinline__ uint64_t wave_match_digit(uint32_t d, uint64_t active)
{
  uint64_t peers = active;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    bool const bit   = (d >> b) & 1u;
    uint64_t const m = ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Adds one digit per lane (kRadix: no digit) into an LDS histogram: one LDS atomic
// when every counted lane has the same digit (skewed data), else one per lane.
__device__ __forceinline__ void wave_count_digit(uint32_t* hist, uint32_t d)
{
  bool const valid   = d < static_cast<uint32_t>(kRadix);
  uint64_t const act = ballot(valid);
  if (act == 0) return;
  int const first   = __builtin_ctzll(act);
  uint32_t const d0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(d), first));
  if (ballot(d == d0) == act) {
    if (lane_id() == first) atomicAdd(&hist[d0], static_cast<uint32_t>(__popcll(act)));
  } else if (valid) {
    atomicAdd(&hist[d], 1u);
  }
}

}  // namespace qpip
