// Device code of the quadrant / bounding-box join shared by the single-tree path
// and the batched (many trees) path: node-box classification and the descent of
// one breadth-first level.
#pragma once

#include "qpip_scan.hpp"

namespace qpip {

constexpr int kMaxJoinLevels = 16;

struct SplitTotal {
  uint64_t leaf;
  uint64_t quad;
  __device__ SplitTotal& operator=(uint64_t packed)
  {
    leaf = packed & 0xFFFFFFFFull;
    quad = packed >> 32;
    return *this;
  }
};

struct JoinMeta {
  SplitTotal level[kMaxJoinLevels];  // per level: leaf pairs emitted, next frontier size
  uint64_t top_pairs;                // T * B
  uint32_t num_top;                  // T = count(level == 0)
  uint32_t max_leaf_offset;
  uint32_t overflow;
  uint32_t bad_index;
  uint64_t need_frontier;
  uint64_t need_output;
};

// z_order.cuh:80-94
__device__ __forceinline__ uint32_t undilate_2(uint32_t x)
{
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0F0F0F0Fu;
  x = (x | (x >> 4)) & 0x00FF00FFu;
  x = (x | (x >> 8)) & 0x0000FFFFu;
  return x;
}

struct TreeView {
  uint32_t const* key;
  uint8_t const* level;
  uint8_t const* internal;
  uint32_t const* length;
  uint32_t const* offset;
  uint64_t m;
};

struct BoxView {
  double const* minx;
  double const* miny;
  double const* maxx;
  double const* maxy;
  uint64_t b;
};

// Geometry of the (single) quadtree: every node shares v_min, scale and depth.
struct Params {
  double vmin_x, vmin_y, scale;
  int max_depth;
  __device__ __forceinline__ void get(uint32_t, double& vx, double& vy, double& sc, int& d) const
  {
    vx = vmin_x;
    vy = vmin_y;
    sc = scale;
    d  = max_depth;
  }
};

// intersection.cuh:99-128: 0 leaf, 1 quad, 2 none. The node extent is
// v_min + key * level_scale, which nvcc contracts to an FMA (default --fmad=true).
template <class Geom>
__device__ __forceinline__ int classify(TreeView const& t, BoxView const& bx, Geom const& g,
                                        uint32_t node, uint32_t bbox)
{
  double vmin_x, vmin_y, scale;
  int max_depth;
  g.get(node, vmin_x, vmin_y, scale, max_depth);
  uint32_t const key = t.key[node];
  int const level    = t.level[node];
  double const key_x = static_cast<double>(undilate_2(key));
  double const key_y = static_cast<double>(undilate_2(key >> 1));
  int const shift    = max_depth - 1 - level;
  double const ls    = scale * static_cast<double>(1 << (shift < 0 ? 0 : shift));
  double const nx0   = fma(key_x + 0, ls, vmin_x);
  double const ny0   = fma(key_y + 0, ls, vmin_y);
  double const nx1   = fma(key_x + 1, ls, vmin_x);
  double const ny1   = fma(key_y + 1, ls, vmin_y);
  if ((nx0 > bx.maxx[bbox]) || (nx1 < bx.minx[bbox]) || (ny0 > bx.maxy[bbox]) ||
      (ny1 < bx.miny[bbox]))
    return 2;
  return t.internal[node] ? 1 : 0;
}

__device__ __forceinline__ uint64_t level_out_base(JoinMeta const* meta, int lvl)
{
  uint64_t b = 0;
  for (int l = 0; l < lvl; ++l) b += meta->level[l].leaf;
  return b;
}

// Level lvl >= 1: expand each quad pair of the previous level to its children
// (traversal.cuh:63-145), classify, and split into leaf output / next frontier.
template <class Geom>
struct DescendOp {
  TreeView t;
  BoxView bx;
  Geom p;
  JoinMeta* meta;
  uint32_t const* in_node;
  uint32_t const* in_bbox;
  uint32_t* out_node;
  uint32_t* out_bbox;
  uint32_t* fr_node;
  uint32_t* fr_bbox;
  uint64_t cap_out, cap_fr;
  int lvl;

  __device__ bool child_range(uint64_t i, uint32_t& first, uint32_t& n) const
  {
    // After a frontier overflow the recorded length exceeds what was stored;
    // the level is rerun with larger buffers, so the excess is skipped here.
    if (i >= cap_fr) return false;
    uint32_t const parent = in_node[i];
    first                 = t.offset[parent];
    n                     = t.length[parent];
    if (static_cast<uint64_t>(first) + n > t.m) {
      atomicOr(&meta->bad_index, 1u);
      return false;
    }
    return true;
  }
  __device__ uint64_t count(uint64_t i) const
  {
    uint32_t first, n;
    if (!child_range(i, first, n)) return 0;
    uint32_t const bbox = in_bbox[i];
    uint64_t acc        = 0;
    for (uint32_t j = 0; j < n; ++j) {
      int const c = classify(t, bx, p, first + j, bbox);
      acc += c == 0 ? 1ull : (c == 1 ? (1ull << 32) : 0ull);
    }
    return acc;
  }
  __device__ void emit(uint64_t i, uint64_t prefix) const
  {
    uint32_t first, n;
    if (!child_range(i, first, n)) return;
    uint32_t const bbox = in_bbox[i];
    uint64_t o          = level_out_base(meta, lvl) + (prefix & 0xFFFFFFFFull);
    uint64_t f          = prefix >> 32;
    for (uint32_t j = 0; j < n; ++j) {
      uint32_t const node = first + j;
      int const c         = classify(t, bx, p, node, bbox);
      if (c == 0) {
        if (o < cap_out) {
          out_node[o] = node;
          out_bbox[o] = bbox;
        } else {
          atomicOr(&meta->overflow, 1u);
        }
        ++o;
      } else if (c == 1) {
        if (f < cap_fr) {
          fr_node[f] = node;
          fr_bbox[f] = bbox;
        } else {
          atomicOr(&meta->overflow, 1u);
        }
        ++f;
      }
    }
  }
};

}  // namespace qpip
