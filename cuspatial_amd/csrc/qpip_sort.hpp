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

// Adds the digits of one wave-wide slot into an LDS histogram, one LDS atomic per
// distinct digit (the lowest lane of each peer group adds the group's size).
__device__ __forceinline__ void wave_hist_add(uint32_t* hist, uint32_t d, bool valid)
{
  uint64_t const active = ballot(valid);
  uint64_t const peers  = wave_match_digit<kRadixBits>(d, active);
  if (valid) {
    int const leader = __builtin_ctzll(peers);
    if (lane_id() == leader) atomicAdd(&hist[d], static_cast<uint32_t>(__popcll(peers)));
  }
}

}  // namespace qpip
