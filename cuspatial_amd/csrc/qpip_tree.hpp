// Device code of the quadtree shared by the single-tree path and the batched
// (many trees) path: Morton keys and the top-down level expansion.
#pragma once

#include "qpip_scan.hpp"

namespace qpip {

// z-order (detail/utility/z_order.cuh:25-77): spread bit i to bit 2i.
__device__ __forceinline__ uint32_t dilate_2(uint32_t x)
{
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

// static_cast<uint16_t> of a non-negative double, NaN -> 0, saturating.
__device__ __forceinline__ uint32_t to_u16(double v)
{
  if (!(v >= 0.0)) return 0u;
  if (v >= 65535.0) return 65535u;
  return static_cast<uint32_t>(v);
}

// First index in [lo, hi) with K[i] >= v.
__device__ __forceinline__ uint32_t lower_bound_u32(uint32_t const* K, uint32_t lo, uint32_t hi,
                                                    uint64_t v)
{
  while (lo < hi) {
    uint32_t const mid = lo + ((hi - lo) >> 1);
    if (static_cast<uint64_t>(K[mid]) < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Key of one point (phase_1.cuh:73-89): outside the box, the last bottom-level key.
__device__ __forceinline__ uint32_t point_key(double x, double y, double min_x, double min_y,
                                              double max_x, double max_y, double scale,
                                              uint32_t outside_key)
{
  if (x < min_x || x > max_x || y < min_y || y > max_y) return outside_key;
  uint32_t const cx = to_u16((x - min_x) / scale);
  uint32_t const cy = to_u16((y - min_y) / scale);
  return (dilate_2(cy) << 1) | dilate_2(cx);
}

// --- level expansion: two launches per level --------------------------------------
// level_split_kernel finds, for every node of level l with more than max_size
// points (l <= d-2: the reference's internal nodes), the boundaries of its four
// child quadrants inside its sorted point range, stores them with the number of
// non-empty quadrants, and sums those per block. level_emit_kernel gives each
// block the sum of the blocks before it, scans its nodes' child counts and appends
// the children in (parent, quadrant) order: level-major and key order within a
// level, the order of reverse_tree_levels.
//
// Node-to-block assignment (both kernels): a level of at most 4 * kLevelBlocks
// nodes is searched one wave per node (block b: nodes 4b .. 4b+3), 21 lanes per
// boundary, each probing one position, so the interval shrinks 22-fold per round
// trip to memory. Larger levels give each block a contiguous chunk and each thread
// one node at a time (three interleaved binary searches).
constexpr int kLevelBlocks = 2048;

// Every kKeySample-th sorted key, written by the sort's last pass (samples[j] =
// K[j * kKeySample]): a search first narrows its range on the samples (an array
// 1/32 the size of the keys, which stays in the caches) and then reads one
// 128-byte line of keys, instead of a line per probe.
constexpr uint32_t kKeySampleShift = 5;
constexpr uint32_t kKeySample      = 1u << kKeySampleShift;

struct LevelArgs {
  uint32_t const* K;  // sorted keys
  uint32_t* nkey;
  uint32_t* nstart;
  uint32_t* ncount;
  uint8_t* nlevel;
  uint32_t* cfirst;
  uint32_t* ccount;
  uint4* split;  // {b1, b2, b3, non-empty quadrants} per node
  uint32_t* partials;
  uint64_t* sizes;     // nodes per level (device)
  uint32_t* overflow;  // set when the node capacity is exceeded
  uint64_t cap;
  int level;
  uint32_t const* samples = nullptr;  // K[j * kKeySample], or null: search K alone
};

// Split rule of one quadtree (every node has the same parameters): a node of
// `level` may have children when level <= d - 2 and it holds more than max_size
// points; its children's key prefixes start at bit 2 (d - 2 - level).
struct UniformLevels {
  int d;
  uint32_t max_size;
  __device__ __forceinline__ bool rule(uint64_t, int level, uint32_t& ms, int& sh,
                                       uint64_t& prefix) const
  {
    ms     = max_size;
    sh     = 2 * (d - 2 - level);
    prefix = 0;
    return level <= d - 2;
  }
  __device__ __forceinline__ void inherit(uint64_t, uint64_t) const {}
};

__device__ __forceinline__ uint64_t level_begin(uint64_t const* sizes, int level)
{
  uint64_t b = 0;
  for (int l = 0; l < level; ++l) b += sizes[l];
  return b;
}

__device__ __forceinline__ bool wave_mode(uint64_t S) { return S <= 4ull * kLevelBlocks; }

__device__ __forceinline__ void node_chunk(uint64_t S, uint64_t b, uint64_t& lo, uint64_t& hi)
{
  if (wave_mode(S)) {
    lo = 4 * b;
    hi = lo + 4;
    if (lo > S) lo = S;
    if (hi > S) hi = S;
  } else {
    chunk_range(S, b, kLevelBlocks, lo, hi);
  }
}

__device__ __forceinline__ uint32_t nonempty(uint32_t s, uint32_t e, uint32_t b1, uint32_t b2,
                                             uint32_t b3)
{
  return (b1 > s) + (b2 > b1) + (b3 > b2) + (e > b3);
}

// Child boundaries b1..b3 of the node (key, [s, e)) with a whole wave; the result
// is returned to every lane. Child q's keys start at prefix | (4 key + q) << sh.
__device__ __forceinline__ void wave_split(uint32_t const* __restrict__ K, uint32_t s, uint32_t e,
                                           uint32_t key, int sh, uint64_t prefix, uint32_t& b1,
                                           uint32_t& b2, uint32_t& b3)
{
  constexpr int P   = 21;  // probes per boundary
  int const lane    = lane_id();
  int const grp     = lane / P;  // boundary 1..3 -> group 0..2; lane 63 idles
  int const pj      = lane - grp * P;
  bool const active = grp < 3;
  uint64_t const target =
    active ? (prefix | ((static_cast<uint64_t>(key) * 4 + static_cast<uint64_t>(grp + 1)) << sh))
           : 0;
  uint64_t const gmask = active ? (((1ull << P) - 1) << (P * grp)) : 0;
  uint32_t lo = s, hi = e;
  while (true) {
    bool const open = active && lo < hi;
    if (ballot(open) == 0) break;
    uint64_t const len = hi - lo;
    uint32_t const pos = lo + static_cast<uint32_t>((static_cast<uint64_t>(pj + 1) * len) / (P + 1));
    bool const below   = open && static_cast<uint64_t>(K[pos]) < target;
    uint64_t const bm  = ballot(below);
    if (open) {
      // probes are monotone: the first `ft` of this group are below the target
      int const ft = __popcll(bm & gmask);
      uint32_t const nlo =
        ft > 0 ? lo + static_cast<uint32_t>((static_cast<uint64_t>(ft) * len) / (P + 1)) + 1 : lo;
      uint32_t const nhi =
        ft < P ? lo + static_cast<uint32_t>((static_cast<uint64_t>(ft + 1) * len) / (P + 1)) : hi;
      lo = nlo;
      hi = nhi;
    }
  }
  b1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lo), 0));
  b2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lo), P));
  b3 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lo), 2 * P));
}

// Child boundaries with one thread: three interleaved binary searches.
__device__ __forceinline__ void thread_split(uint32_t const* __restrict__ K, uint32_t s,
                                             uint32_t e, uint32_t key, int sh, uint64_t prefix,
                                             uint32_t& b1, uint32_t& b2, uint32_t& b3)
{
  uint64_t t[3];
  uint32_t lo[3], hi[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    t[q]  = prefix | ((static_cast<uint64_t>(key) * 4 + static_cast<uint64_t>(q + 1)) << sh);
    lo[q] = s;
    hi[q] = e;
  }
  while (lo[0] < hi[0] || lo[1] < hi[1] || lo[2] < hi[2]) {
    uint32_t mid[3], v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      mid[q] = lo[q] + ((hi[q] - lo[q]) >> 1);
      v[q]   = lo[q] < hi[q] ? K[mid[q]] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (lo[q] < hi[q]) {
        if (static_cast<uint64_t>(v[q]) < t[q])
          lo[q] = mid[q] + 1;
        else
          hi[q] = mid[q];
      }
    }
  }
  b1 = lo[0];
  b2 = lo[1];
  b3 = lo[2];
}

// thread_split over the key samples: three interleaved lower bounds on the
// samples inside [s, e), then on the (at most 32) keys each one leaves.
// Same results as thread_split: with a = ceil(s / 32), b = floor((e - 1) / 32)
// and j* the first sample in [a, b] at or above the target, the first key at or
// above the target lies in [s, 32a] (j* = a), (32(j* - 1), 32 j*] (a < j* <= b)
// or (32 b, e] (no sample above it).
__device__ __forceinline__ void sampled_split(uint32_t const* __restrict__ K,
                                              uint32_t const* __restrict__ S, uint32_t s,
                                              uint32_t e, uint32_t key, int sh, uint64_t prefix,
                                              uint32_t& b1, uint32_t& b2, uint32_t& b3)
{
  uint32_t const a = (s + kKeySample - 1) >> kKeySampleShift;
  uint32_t const b = (e - 1) >> kKeySampleShift;  // e > s
  uint64_t t[3];
  uint32_t lo[3], hi[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    t[q]  = prefix | ((static_cast<uint64_t>(key) * 4 + static_cast<uint64_t>(q + 1)) << sh);
    lo[q] = a;
    hi[q] = a <= b ? b + 1 : a;  // no sample inside: the window is [s, e)
  }
  while (lo[0] < hi[0] || lo[1] < hi[1] || lo[2] < hi[2]) {
    uint32_t mid[3], v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      mid[q] = lo[q] + ((hi[q] - lo[q]) >> 1);
      v[q]   = lo[q] < hi[q] ? S[mid[q]] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (lo[q] < hi[q]) {
        if (static_cast<uint64_t>(v[q]) < t[q])
          lo[q] = mid[q] + 1;
        else
          hi[q] = mid[q];
      }
    }
  }
  // lo[q] = j*; the window of keys that holds the answer
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    uint32_t const j = lo[q];
    uint32_t wlo, whi;
    if (a > b) {
      wlo = s, whi = e;
    } else if (j == a) {
      wlo = s, whi = (a << kKeySampleShift) + 1;
    } else if (j <= b) {
      wlo = ((j - 1) << kKeySampleShift) + 1, whi = (j << kKeySampleShift) + 1;
    } else {
      wlo = (b << kKeySampleShift) + 1, whi = e;
    }
    lo[q] = wlo;
    hi[q] = whi;
  }
  while (lo[0] < hi[0] || lo[1] < hi[1] || lo[2] < hi[2]) {
    uint32_t mid[3], v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      mid[q] = lo[q] + ((hi[q] - lo[q]) >> 1);
      v[q]   = lo[q] < hi[q] ? K[mid[q]] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (lo[q] < hi[q]) {
        if (static_cast<uint64_t>(v[q]) < t[q])
          lo[q] = mid[q] + 1;
        else
          hi[q] = mid[q];
      }
    }
  }
  b1 = lo[0];
  b2 = lo[1];
  b3 = lo[2];
}

template <class Rule>
__global__ void __launch_bounds__(256) level_split_kernel(LevelArgs a, Rule rule)
{
  __shared__ uint32_t lds[256 / kWave + 1];
  uint64_t const S     = a.sizes[a.level];
  uint64_t const begin = level_begin(a.sizes, a.level);
  uint64_t lo, hi;
  node_chunk(S, blockIdx.x, lo, hi);
  uint32_t acc = 0;
  if (wave_mode(S)) {
    uint64_t const i    = lo + threadIdx.x / kWave;
    uint64_t const node = begin + i;
    if (i < hi && node < a.cap) {
      uint32_t const c = a.ncount[node];
      uint4 sp         = make_uint4(0, 0, 0, 0);
      uint32_t ms;
      int sh;
      uint64_t prefix;
      if (rule.rule(node, a.level, ms, sh, prefix) && c > ms) {
        uint32_t const s = a.nstart[node];
        uint32_t b1, b2, b3;
        wave_split(a.K, s, s + c, a.nkey[node], sh, prefix, b1, b2, b3);
        sp = make_uint4(b1, b2, b3, nonempty(s, s + c, b1, b2, b3));
      }
      if (lane_id() == 0) {
        a.split[node] = sp;
        acc           = sp.w;
      }
    }
  } else {
    for (uint64_t i = lo + threadIdx.x; i < hi; i += 256) {
      uint64_t const node = begin + i;
      if (node >= a.cap) break;  // only after an overflow (reported to the host)
      uint32_t const c = a.ncount[node];
      uint4 sp         = make_uint4(0, 0, 0, 0);
      uint32_t ms;
      int sh;
      uint64_t prefix;
      if (rule.rule(node, a.level, ms, sh, prefix) && c > ms) {
        uint32_t const s = a.nstart[node];
        uint32_t b1, b2, b3;
        if (a.samples)
          sampled_split(a.K, a.samples, s, s + c, a.nkey[node], sh, prefix, b1, b2, b3);
        else
          thread_split(a.K, s, s + c, a.nkey[node], sh, prefix, b1, b2, b3);
        sp = make_uint4(b1, b2, b3, nonempty(s, s + c, b1, b2, b3));
      }
      a.split[node] = sp;
      acc += sp.w;
    }
  }
  uint32_t tot;
  block_excl_scan<256>(acc, lds, &tot);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = tot;
}

template <class Rule>
__global__ void __launch_bounds__(256) level_emit_kernel(LevelArgs a, Rule rule)
{
  __shared__ uint32_t lds[256 / kWave + 1];
  uint64_t const S     = a.sizes[a.level];
  uint64_t const begin = level_begin(a.sizes, a.level);
  // this block's base (children of earlier blocks) and the level's total
  uint32_t before = 0, all = 0;
  for (int i = threadIdx.x; i < kLevelBlocks; i += 256) {
    uint32_t const p = a.partials[i];
    all += p;
    if (i < static_cast<int>(blockIdx.x)) before += p;
  }
  uint32_t base, total;
  block_excl_scan<256>(before, lds, &base);
  block_excl_scan<256>(all, lds, &total);
  if (blockIdx.x == 0 && threadIdx.x == 0) a.sizes[a.level + 1] = total;
  uint64_t lo, hi;
  node_chunk(S, blockIdx.x, lo, hi);
  uint64_t const next0 = begin + S + base;
  uint64_t run         = 0;
  for (uint64_t t0 = lo; t0 < hi; t0 += 256) {
    uint64_t const i    = t0 + threadIdx.x;
    uint64_t const node = begin + i;
    bool const live     = i < hi && node < a.cap;
    uint4 const sp      = live ? a.split[node] : make_uint4(0, 0, 0, 0);
    uint32_t tile_total;
    uint32_t const ex = block_excl_scan<256>(sp.w, lds, &tile_total);
    if (live) {
      uint64_t const next = next0 + run + ex;  // first child slot
      a.cfirst[node]      = static_cast<uint32_t>(next);
      a.ccount[node]      = sp.w;
      if (sp.w > 0) {
        if (next + sp.w > a.cap) {
          atomicOr(a.overflow, 1u);
        } else {
          uint32_t const k = a.nkey[node];
          uint32_t const s = a.nstart[node];
          uint32_t const e = s + a.ncount[node];
          uint32_t const b[5] = {s, sp.x, sp.y, sp.z, e};
          uint64_t slot       = next;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (b[q + 1] > b[q]) {
              a.nkey[slot]   = k * 4 + q;
              a.nstart[slot] = b[q];
              a.ncount[slot] = b[q + 1] - b[q];
              a.nlevel[slot] = static_cast<uint8_t>(a.level + 1);
              rule.inherit(node, slot);
              ++slot;
            }
          }
        }
      }
    }
    run += tile_total;
  }
}


}  // namespace qpip
