// Effective-edge table of polygon rings, shared by the quadtree point-in-polygon
// test and the grid locator: the point-independent part of
// detail/algorithm/is_point_in_polygon.cuh:46-100 (which ring edges the test
// visits, in which order), built once per polygon set.
#pragma once

#include "qpip_internal.hpp"

namespace qpip {

namespace {

// Hot part of an effective edge (read by every point): 32 bytes, one scalar load.
struct alignas(32) Edge {
  double ax, ay, run, rise;
};
// Cold part: the edge's `b` vertex (ring start y0 and the on-edge x range).
struct alignas(16) EdgeB {
  double bx, by;
};

// float_equal (detail/utility/floating_point.cuh:97-129): |biased(a) - biased(b)| <= 4.
__device__ __forceinline__ bool float_equal(double a, double b)
{
  if (isnan(a) || isnan(b)) return false;
  uint64_t const sign = 1ull << 63;
  uint64_t ua         = static_cast<uint64_t>(__double_as_longlong(a));
  uint64_t ub         = static_cast<uint64_t>(__double_as_longlong(b));
  ua                  = (ua & sign) ? (~ua + 1) : (ua | sign);
  ub                  = (ub & sign) ? (~ub + 1) : (ub | sign);
  return ua >= ub ? (ua - ub) <= 4 : (ub - ua) <= 4;
}



// --- 1. edge table: one wave per ring ----------------------------------------------
// The reference walks a ring from its last vertex and skips a vertex whose edge to
// the current `b` is zero within 4 ULP, without moving `b` (is_point_in_polygon.cuh:65).
// In the common case `b` is simply the previous vertex: every vertex v > first is
// kept when compared with v - 1, and the first vertex is kept, or skipped while
// bitwise equal to the last one (a closed ring). Then the lanes build the edges in
// parallel (ballot + popcount give each kept vertex its slot). Any other ring is
// walked sequentially by lane 0, exactly as the reference does.
// One ring, by one wave (build_edges_kernel below strides its waves over rings).
__device__ __forceinline__ void build_ring_edges(uint32_t const* __restrict__ ring_offsets,
                                                 int64_t r,
                                                 double const* __restrict__ vx,
                                                 double const* __restrict__ vy,
                                                 int64_t nv,
                                                 Edge* __restrict__ edges,
                                                 EdgeB* __restrict__ edges_b,
                                                 uint32_t* __restrict__ ring_edges)
{
  int const lane = lane_id();
  int64_t vb     = ring_offsets[r];
  int64_t ve     = ring_offsets[r + 1];
  if (vb > nv) vb = nv;
  if (ve > nv) ve = nv;
  if (ve <= vb) {
    if (lane == 0) ring_edges[r] = 0;
    return;
  }
  double const lx = vx[ve - 1], ly = vy[ve - 1];
  // pass 1: is the pairwise rule the reference's walk?
  bool simple = true;
  for (int64_t v0 = vb; v0 < ve; v0 += kWave) {
    int64_t const v = v0 + lane;
    bool bad        = false;
    if (v < ve) {
      double const ax = vx[v], ay = vy[v];
      double const bx = v == vb ? lx : vx[v - 1];
      double const by = v == vb ? ly : vy[v - 1];
      bool const skip = float_equal(bx - ax, 0.0) && float_equal(by - ay, 0.0);
      if (skip) bad = v != vb || !(ax == lx && ay == ly &&
                                    __double_as_longlong(ax) == __double_as_longlong(lx) &&
                                    __double_as_longlong(ay) == __double_as_longlong(ly));
    }
    if (ballot(bad)) simple = false;
  }
  if (simple) {
    uint32_t base = 0;
    for (int64_t v0 = vb; v0 < ve; v0 += kWave) {
      int64_t const v = v0 + lane;
      bool keep       = false;
      double ax = 0, ay = 0, bx = 0, by = 0;
      if (v < ve) {
        ax   = vx[v];
        ay   = vy[v];
        bx   = v == vb ? lx : vx[v - 1];
        by   = v == vb ? ly : vy[v - 1];
        keep = !(float_equal(bx - ax, 0.0) && float_equal(by - ay, 0.0));
      }
      uint64_t const km = ballot(keep);
      if (keep) {
        uint32_t const slot = base + static_cast<uint32_t>(__popcll(km & lanemask_lt()));
        edges[vb + slot]    = Edge{ax, ay, bx - ax, by - ay};
        edges_b[vb + slot]  = EdgeB{bx, by};
      }
      base += static_cast<uint32_t>(__popcll(km));
    }
    if (lane == 0) ring_edges[r] = base;
    return;
  }
  if (lane != 0) return;
  uint32_t ne = 0;
  double bx = lx, by = ly;
  for (int64_t v = vb; v < ve; ++v) {
    double const ax = vx[v], ay = vy[v];
    double const run = bx - ax, rise = by - ay;
    if (float_equal(run, 0.0) && float_equal(rise, 0.0)) continue;
    edges[vb + ne]   = Edge{ax, ay, run, rise};
    edges_b[vb + ne] = EdgeB{bx, by};
    ++ne;
    bx = ax;
    by = ay;
  }
  ring_edges[r] = ne;
}

__global__ void __launch_bounds__(256)
  build_edges_kernel(uint32_t const* __restrict__ ring_offsets,
                     int64_t num_rings,
                     double const* __restrict__ vx,
                     double const* __restrict__ vy,
                     int64_t nv,
                     Edge* __restrict__ edges,
                     EdgeB* __restrict__ edges_b,
                     uint32_t* __restrict__ ring_edges)
{
  int64_t const waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
       r < num_rings; r += waves)
    build_ring_edges(ring_offsets, r, vx, vy, nv, edges, edges_b, ring_edges);
}

}  // namespace

}  // namespace qpip
