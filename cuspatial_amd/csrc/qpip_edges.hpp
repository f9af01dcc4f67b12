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
// Vertex box of a ring or polygon for the test's point filter, one scalar load.
// Empty (min > max): no vertices. Infinite: the filter may not be used (a
// coordinate that is not finite or above 2^400 in magnitude, or an edge rise
// that is non-zero and below 2^-600 in magnitude).
struct alignas(32) PolyBox {
  double minx, maxx, miny, maxy;
};
constexpr double kBoxCoordMax = 0x1p400;
constexpr double kBoxRiseMin  = 0x1p-600;

// An effective edge in the form the PIP test's fast loop reads it (from LDS):
// the line t = px * rise' - py * run' + c' of the edge scaled by 2^-k, the power
// of two just above n = |run| + |rise| (an exact scaling), with
// c' = ay * run' - ax * rise' (one product, one fma), negated where rise < 0 so
// that t < 0 means "the crossing toggles" wherever the edge straddles the point
// (there y1 = (ay > by) = (rise < 0)). An edge the test's certain-sign bound does
// not cover (n outside [2^-70, 2^400], a coordinate that is not finite, or a
// scaling that would lose bits) gets c' = NaN: |t| > tol is then false, so every
// point is retested exactly on it.
struct alignas(32) LineEdge {
  double ay, run, rise, c;
};

__device__ __forceinline__ LineEdge line_of(double ax, double ay, double run, double rise)
{
  double const n = fabs(run) + fabs(rise);
  if (isfinite(ax) && isfinite(ay) && n >= 0x1p-70 && n <= 0x1p400) {
    int const k     = ilogb(n) + 1;
    double const rs = ldexp(run, -k), ss = ldexp(rise, -k);
    if (ldexp(rs, k) == run && ldexp(ss, k) == rise) {
      double const c = fma(ay, rs, -(ax * ss));
      return ss < 0.0 ? LineEdge{ay, -rs, -ss, -c} : LineEdge{ay, rs, ss, c};
    }
  }
  return LineEdge{ay, 0.0, 0.0, __builtin_nan("")};
}

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
// A vertex the filter box can vouch for: finite and at most 2^400 in magnitude
// (false for NaN).
__device__ __forceinline__ bool box_coord_ok(double v) { return fabs(v) <= kBoxCoordMax; }
// An edge rise the filter's proof holds for: zero (the edge never straddles a
// point) or at least 2^-600 in magnitude.
__device__ __forceinline__ bool box_rise_ok(double rise)
{
  return rise == 0.0 || fabs(rise) >= kBoxRiseMin;
}

// Ring `r`'s largest |coordinate| (NaN ignored; infinite with an infinite
// vertex) and box, from per-lane partial extrema and flags (all lanes take part).
__device__ __forceinline__ void finish_ring(double* ring_amax, PolyBox* ring_box, int64_t r,
                                            double mnx, double mxx, double mny, double mxy,
                                            double amax, bool ok)
{
  if (!ring_amax && !ring_box) return;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mnx  = fmin(mnx, __shfl_xor(mnx, o, kWave));
    mxx  = fmax(mxx, __shfl_xor(mxx, o, kWave));
    mny  = fmin(mny, __shfl_xor(mny, o, kWave));
    mxy  = fmax(mxy, __shfl_xor(mxy, o, kWave));
    amax = fmax(amax, __shfl_xor(amax, o, kWave));
  }
  bool const all_ok = ballot(!ok) == 0;
  if (lane_id() == 0) {
    double const inf = __builtin_inf();
    if (ring_amax) ring_amax[r] = amax;
    if (ring_box) ring_box[r] = all_ok ? PolyBox{mnx, mxx, mny, mxy} : PolyBox{-inf, inf, -inf, inf};
  }
}

// One ring, by one wave (build_edges_kernel below strides its waves over rings).
// With `ring_box`, also the ring's vertex box (PolyBox above).
__device__ __forceinline__ void build_ring_edges(uint32_t const* __restrict__ ring_offsets,
                                                 int64_t r,
                                                 double const* __restrict__ vx,
                                                 double const* __restrict__ vy,
                                                 int64_t nv,
                                                 Edge* __restrict__ edges,
                                                 EdgeB* __restrict__ edges_b,
                                                 uint32_t* __restrict__ ring_edges,
                                                 LineEdge* __restrict__ lines,
                                                 double* __restrict__ ring_amax,
                                                 PolyBox* __restrict__ ring_box)
{
  int const lane   = lane_id();
  double const inf = __builtin_inf();
  int64_t vb       = ring_offsets[r];
  int64_t ve       = ring_offsets[r + 1];
  if (vb > nv) vb = nv;
  if (ve > nv) ve = nv;
  if (ve <= vb) {
    if (lane == 0) {
      ring_edges[r] = 0;
      if (ring_amax) ring_amax[r] = 0.0;
      if (ring_box) ring_box[r] = PolyBox{inf, -inf, inf, -inf};
    }
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
  // the box covers every vertex of the ring (a superset of the effective edges' ends)
  double mnx = inf, mxx = -inf, mny = inf, mxy = -inf, amax = 0.0;
  bool ok = true;
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
        mnx  = fmin(mnx, ax);
        mxx  = fmax(mxx, ax);
        mny  = fmin(mny, ay);
        mxy  = fmax(mxy, ay);
        amax = fmax(amax, fmax(fabs(ax), fabs(ay)));
        ok   = ok && box_coord_ok(ax) && box_coord_ok(ay) && (!keep || box_rise_ok(by - ay));
      }
      uint64_t const km = ballot(keep);
      if (keep) {
        uint32_t const slot = base + static_cast<uint32_t>(__popcll(km & lanemask_lt()));
        if (edges) edges[vb + slot] = Edge{ax, ay, bx - ax, by - ay};
        if (edges_b) edges_b[vb + slot] = EdgeB{bx, by};
        if (lines) lines[vb + slot] = line_of(ax, ay, bx - ax, by - ay);
      }
      base += static_cast<uint32_t>(__popcll(km));
    }
    if (lane == 0) ring_edges[r] = base;
    finish_ring(ring_amax, ring_box, r, mnx, mxx, mny, mxy, amax, ok);
    return;
  }
  if (lane == 0) {
    uint32_t ne = 0;
    double bx = lx, by = ly;
    for (int64_t v = vb; v < ve; ++v) {
      double const ax = vx[v], ay = vy[v];
      mnx  = fmin(mnx, ax);
      mxx  = fmax(mxx, ax);
      mny  = fmin(mny, ay);
      mxy  = fmax(mxy, ay);
      amax = fmax(amax, fmax(fabs(ax), fabs(ay)));
      ok   = ok && box_coord_ok(ax) && box_coord_ok(ay);
      double const run = bx - ax, rise = by - ay;
      if (float_equal(run, 0.0) && float_equal(rise, 0.0)) continue;
      ok               = ok && box_rise_ok(rise);
      if (edges) edges[vb + ne] = Edge{ax, ay, run, rise};
      if (edges_b) edges_b[vb + ne] = EdgeB{bx, by};
      if (lines) lines[vb + ne] = line_of(ax, ay, run, rise);
      ++ne;
      bx = ax;
      by = ay;
    }
    ring_edges[r] = ne;
  }
  finish_ring(ring_amax, ring_box, r, mnx, mxx, mny, mxy, amax, ok);
}

__global__ void __launch_bounds__(256)
  build_edges_kernel(uint32_t const* __restrict__ ring_offsets,
                     int64_t num_rings,
                     double const* __restrict__ vx,
                     double const* __restrict__ vy,
                     int64_t nv,
                     Edge* __restrict__ edges,
                     EdgeB* __restrict__ edges_b,
                     uint32_t* __restrict__ ring_edges,
                     LineEdge* __restrict__ lines,   // null: no line forms
                     double* __restrict__ ring_amax,  // null: no ring magnitudes
                     PolyBox* __restrict__ ring_box)  // null: no boxes
{
  int64_t const waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
       r < num_rings; r += waves)
    build_ring_edges(ring_offsets, r, vx, vy, nv, edges, edges_b, ring_edges, lines, ring_amax,
                     ring_box);
}

// Polygon boxes: the union of each polygon's ring boxes (ring offsets past the
// ring count are clamped, as the test clamps them).
__global__ void __launch_bounds__(256)
  poly_box_kernel(uint32_t const* __restrict__ poly_offsets, uint64_t num_polys,
                  uint64_t num_rings, PolyBox const* __restrict__ ring_box,
                  PolyBox* __restrict__ poly_box)
{
  double const inf = __builtin_inf();
  for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < num_polys;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint64_t r0 = poly_offsets[p], r1 = poly_offsets[p + 1];
    if (r1 > num_rings) r1 = num_rings;
    PolyBox b{inf, -inf, inf, -inf};
    for (uint64_t r = r0; r < r1; ++r) {
      PolyBox const q = ring_box[r];
      b.minx = fmin(b.minx, q.minx);
      b.maxx = fmax(b.maxx, q.maxx);
      b.miny = fmin(b.miny, q.miny);
      b.maxy = fmax(b.maxy, q.maxy);
    }
    poly_box[p] = b;
  }
}

}  // namespace

}  // namespace qpip
