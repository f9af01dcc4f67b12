// Device helpers of the LSD radix sort shared with the fused keygen kernel.
#pragma once

#include "qpip_internal.hpp"

namespace qpip {

// Lanes of the wave whose digit equals this lane's digit (restricted to `active`).
// One ballot per digit bit: peers = AND over bits of (ballot(bit) or its complement).
// The complement is taken by XOR with a per-lane all-ones/zero word (the sign-
// extended inverted bit), so each bit costs one signed bit-field extract, one
// compare and one v_bitop3 per 32-bit half computing acc & (ballot ^ flip).
// v_bitop3's table is indexed by (src0 << 2 | src1 << 1 | src2): src0 = 0xF0,
// src1 = 0xCC, src2 = 0xAA, so acc & (ballot ^ flip) = 0xF0 & (0xCC ^ 0xAA) = 0x60.
template <int NB>
__device__ __forceinline__ uint64_t wave_match_digit(uint32_t d, uint64_t active)
{
  constexpr uint32_t kAndXor = 0x60;
  uint32_t lo               = static_cast<uint32_t>(active);
  uint32_t hi               = static_cast<uint32_t>(active >> 32);
  int32_t const nd          = static_cast<int32_t>(~d);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    // all ones where this lane's bit b is 0, zero where it is 1
    uint32_t const flip = static_cast<uint32_t>(__builtin_amdgcn_sbfe(nd, b, 1));
    uint64_t const m    = ballot(flip == 0u);  // lanes whose bit b is 1
    lo = __builtin_amdgcn_bitop3_b32(lo, static_cast<uint32_t>(m), flip, kAndXor);
    hi = __builtin_amdgcn_bitop3_b32(hi, static_cast<uint32_t>(m >> 32), flip, kAndXor);
  }
  return (static_cast<uint64_t>(hi) << 32) | lo;
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
