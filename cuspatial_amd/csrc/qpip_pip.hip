// quadtree_point_in_polygon on gfx950.
//
// Reference: cpp/src/join/quadtree_point_in_polygon.cu:143-191 (column API) and
// detail/join/quadtree_point_in_polygon.cuh:104-235: every point of every
// (polygon, leaf) pair is tested with is_point_in_polygon
// (detail/algorithm/is_point_in_polygon.cuh:46-100) and the hits are compacted
// in pair order, then point order within the leaf.
//
// Design:
//   1. Edge table. The crossing test walks each ring from the last vertex, skips
//      vertices whose edge to the current `b` is zero within 4 ULP (without moving
//      `b`, :65), and tests every other edge (a, b). Which edges are skipped does
//      not depend on the test point, so each ring is walked once here into a list
//      of effective edges, each stored as the scaled line form the fast test reads
//      (LineEdge). The per-point test is then order-independent: inside = parity
//      of crossings over all effective edges, forced to false if any edge is an
//      on-edge hit (the reference breaks out and returns false at the first such
//      edge).
//   2. Chunks. Pair k covers ceil(length / 64) chunks of 64 consecutive point
//      positions; a scan gives each pair its first chunk, and a map gives each
//      chunk its pair (get_quad_and_local_point_indices.cuh:26-45 searches for it).
//   3. Test, in one of two orders. Leaf order (pip_test_kernel): one wave per
//      segment of a leaf's pairs (they are adjacent in pair order), up to 128
//      points (2 chunks) at a time. Polygon order (pip_test_pm_kernel): the
//      pairs counting-sorted by polygon, one wave per unit of up to pm_group of
//      one polygon's pairs, with the ring's edges shared by the unit's leaves;
//      a flush of at most 32 points runs as several lane groups that take
//      different edges of the same batch. Either way the lanes read consecutive
//      point_indices entries (coalesced) and gather the coordinates once; points
//      outside a polygon's vertex box are dropped and the rest packed into fewer
//      lanes; each polygon's edges are wave-uniform and are staged through the
//      wave's LDS slot in coalesced batches. A certain-sign form of the crossing
//      test decides almost every (point, edge); a point with an undecided edge
//      is retested by walking the ring's vertices with the reference's own loop
//      and FP64 arithmetic. The hits of a chunk end up as the bits of its 64-bit
//      mask (a wave's ballot in leaf order, one atomic OR per hit in polygon
//      order, where one flush packs the points of several of the unit's chunks).
//   4. Scan of the masks' popcounts -> output positions; one host sync for the
//      total.
//   5. Write: each wave expands the ballots of 64 chunks into (polygon, point
//      position) pairs.
//   No candidate-sized buffer is ever allocated, so the reference's out-of-memory
//   fallback (count, then allocate; :220-234) is the only path here.
#include <algorithm>
#include <mutex>

#include "qpip_pip.hpp"
#include "qpip_scan.hpp"

namespace qpip {

namespace {

// --- 2. chunk enumeration ---------------------------------------------------------
struct ChunkOp {
  uint32_t const* pair_quad;
  uint32_t const* length;
  uint64_t m;
  uint32_t* chunk_start;
  uint32_t* bad;
  __device__ uint32_t count(uint64_t k) const
  {
    uint32_t const q = pair_quad[k];
    if (q >= m) {
      atomicOr(bad, 1u);
      return 0;
    }
    return (length[q] + kWave - 1) / kWave;
  }
  // Starts past 2^32 - 1 saturate: the test kernel's capacity check rejects them.
  __device__ void emit(uint64_t k, uint64_t prefix) const
  {
    chunk_start[k] = prefix < 0xFFFFFFFFull ? static_cast<uint32_t>(prefix) : 0xFFFFFFFFu;
  }
};

struct PipArgs {
  uint32_t const* pair_poly;
  uint32_t const* pair_quad;
  uint64_t num_pairs;
  uint32_t const* length;
  uint32_t const* offset;
  uint32_t const* point_indices;
  uint64_t num_points;
  double const* x;
  double const* y;
  uint32_t const* poly_offsets;
  uint64_t num_polys;
  uint64_t num_rings;
  uint32_t const* ring_offsets;
  uint32_t const* ring_edges;
  LineEdge const* lines;   // per edge: the fast loop's form
  double const* ring_amax;  // per ring: largest |vertex coordinate|
  uint32_t const* chunk_start;
  uint64_t num_chunks;
  uint64_t num_nodes;
  double const* xy;  // interleaved x, y pairs instead of x / y (null: two arrays)
  bool xy_vec;       // xy is 16-byte aligned
  uint32_t pairs_per_wave;  // longest segment of a leaf's pairs one wave takes
  uint64_t const* num_pairs_dev;  // non-null: the pair count lives on the device
  uint64_t chunk_cap;             // entries of masks / chunk_info
  // non-null: point_indices hold indices local to each node's job, and
  // point_base[node] is the job's first point (the batched forest)
  uint32_t const* point_base = nullptr;
  // non-null: each polygon's vertex box, for the test's point filter
  PolyBox const* poly_box = nullptr;
  // polygon-major order (pip_test_pm_kernel): the valid pairs grouped by
  // polygon, polygon p's at pm_order[pm_start[p] .. + pm_count[p]), each as the
  // record the test needs (PmPair); a unit is up to pm_group of one polygon's
  // pairs, *pm_num_units of them, listed by their first position in pm_order
  uint4 const* pm_order = nullptr;
  uint32_t const* pm_start = nullptr;
  uint32_t const* pm_count = nullptr;
  uint4 const* pm_units = nullptr;  // per unit: first order position, polygon, pairs
  uint64_t const* pm_num_units = nullptr;
  // the polygons' vertices (num_vertices of them): the exact retest walks them
  double const* poly_x  = nullptr;
  double const* poly_y  = nullptr;
  uint64_t num_vertices = 0;
  uint32_t pm_group        = 0;
};

// Segments of at most max(m, pairs >> s) pairs (m = 2, s = 24 by default; both
// are tuning parameters): short segments spread the pairs over many waves, so
// that enough waves hide the latency of the point gathers and edge loads; a
// large s keeps segments at m pairs up to 2^25 pairs, and past that bounds the
// number of waves (each segment re-gathers its leaf's points).
inline uint32_t pairs_per_wave(uint64_t num_pairs)
{
  uint64_t const m = static_cast<uint64_t>(tuning(Tuning::PipMinPairsPerWave));
  uint64_t const g = num_pairs >> tuning(Tuning::PipPairsShift);
  uint64_t const v = g < m ? m : g;
  return static_cast<uint32_t>(v > (1u << 30) ? (1u << 30) : v);
}

// --- 3. test ------------------------------------------------------------------------
// One wave per segment of at most pairs_per_wave pairs that share a leaf (the join
// orders pairs by leaf offset, so such pairs are adjacent; segments start at a run's
// first pair and at every multiple of pairs_per_wave inside a run): the segment's
// first wave gathers the leaf's points up to kSlots chunks (128 points) at a time
// and tests them against every polygon of the segment; the other waves return at
// once. Masks are stored per (pair, chunk) in pair order: chunk j of pair k is
// chunk_start[k] + j.
//
// Certain-sign filter. The reference decides each (point, edge) by the sign of
// lhs - rhs, lhs = (px - ax) * rise and rhs = run * (py - ay), rounded as written,
// and calls the point on the edge when the two are within 4 ULP. Here each edge is
// evaluated in its line form (LineEdge, built with the edge table): the edge
// scaled exactly by 2^-k, the power of two just above n = |run| + |rise|, so that
// n' = |run'| + |rise'| <= 1, with c' = ay * run' - ax * rise' (one product, one
// fma) and t = px * rise' - py * run' + c' (two fmas per point). t decides only
// when it is far from zero. With u = 2^-53, s = (largest |px|, |py| of the wave's
// points) + (largest |ax|, |ay| of the ring), L - R = (px - ax) rise - run (py - ay)
// and L' - R' = (L - R) / 2^k (exact):
//   |t - (L' - R')| <= 7.1 u s n'  (c', the inner and the outer fma),
//   |lhs - L| + |rhs - R| <= 2.01 u s n  (two roundings each),
//   4 ULP of lhs or rhs <= 8.01 u s n, plus 2^-1070 where values are subnormal,
// and dividing the last two by 2^k >= n (>= 2^-69) gives at most 10.02 u s + 2^-1001.
// So |t| > tol = 2^-48 * s + 2^-1000 (over 31.9 u s after rounding) gives
// sign(lhs - rhs) = sign(t) and rules out the 4-ULP collinearity; every other
// (point, edge), NaN and infinity included (their comparisons are false), marks the
// point "suspect". Suspect points of the polygon are retested with the reference's
// arithmetic in a second, rare pass. s is capped at 2^400 (beyond, tol is infinite
// and every point is suspect) and n at [2^-70, 2^400] (outside, the line form's
// c' is NaN), so no product overflows.
//
// Lane state is kept as wave-uniform 64-bit masks (one bit per lane and point
// slot), so the loop's bookkeeping is a handful of scalar ops per edge.
// Ring staging: a wave copies up to S line forms of the ring into its own LDS
// slot with one coalesced vector load per lane (one memory latency per batch),
// then walks them with broadcast LDS reads. Walking them from global memory took
// a dependent scalar load (and its latency) every two edges, which dominated when
// the edge table is not in L2: a small join right after the table is built, or
// the batched join's multi-GB table.
// The depth S (edges per batch, 64 or 128) and the waves per workgroup are runtime
// tuning parameters (pip_stage_edges, pip_waves_per_block).
constexpr int kStageEdges = kWave;  // the default depth, and the bitmask kernel's

// A wave's staged batch of line forms (LineEdge, built with the edge table).
template <int S>
struct PipStage {
  LineEdge e[S];
};

// Point slots a wave holds at a time: lane l of slot u is point u * 64 + l of
// the wave's current group (two slots keep the per-slot masks in few scalar
// registers at eight waves per SIMD).
constexpr int kSlots = 2;

// Largest |coordinate| of the wave's points in slots [0, nu) (NaN ignored).
__device__ __forceinline__ double wave_max_abs(double const (&px)[kSlots],
                                               double const (&py)[kSlots], uint32_t nu)
{
  double m = fmax(fabs(px[0]), fabs(py[0]));
#pragma unroll
  for (int u = 1; u < kSlots; ++u)
    if (static_cast<uint32_t>(u) < nu) m = fmax(m, fmax(fabs(px[u]), fabs(py[u])));
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
  return m;
}

// Lane groups (lg > 0, one slot only): the wave holds G = 2^lg copies of q <= 64 / G
// points, lane l holding point l % (64 / G) of copy g = l / (64 / G), and step e
// gives copy g the ring's edge e * G + g: the G copies read G adjacent line forms.
// A crossing toggles on the two ends of its own edge only: y1 from its first
// vertex, y0 from the vertex before it, which is the previous edge's first vertex
// and so copy g - 1's y1 of the same step (64 / G lanes down the ballot), or, for
// copy 0, the last copy's y1 of the step before. The ring's parity is the XOR of
// the copies' parities (the suspect marks their OR), so a flush of few points
// takes G times fewer edge steps. The ring's last batch is padded to a multiple
// of G with edges that neither toggle (their vertex repeats the ring's last one)
// nor mark (t = +inf for finite points; for a point that is not finite t is NaN,
// a mark, as on every real edge).
// The reference's test of one point against rings [r0, r1) of a polygon
// (is_point_in_polygon.cuh:46-100), on the vertices: a ring's first edge ends at
// its last vertex, an edge with run and rise both within 4 ULP of 0 is skipped
// without moving `b`, a collinear point within an edge's x range is on the edge.
// Vertex ranges are clamped to the vertex count, as the edge table's builder
// clamps them. Returns inside and not on an edge.
__device__ __forceinline__ bool exact_in_polygon(PipArgs const& a, uint32_t r0, uint32_t r1,
                                                 double px, double py)
{
  bool inside = false, on_edge = false;
  for (uint32_t r = r0; r < r1; ++r) {
    uint64_t vb = a.ring_offsets[r], ve = a.ring_offsets[r + 1];
    if (vb > a.num_vertices) vb = a.num_vertices;
    if (ve > a.num_vertices) ve = a.num_vertices;
    if (ve <= vb) continue;
    double bx = a.poly_x[ve - 1], by = a.poly_y[ve - 1];
    bool y0 = by > py;
    for (uint64_t v = vb; v < ve; ++v) {
      double const ax = a.poly_x[v], ay = a.poly_y[v];
      double const run = bx - ax, rise = by - ay;
      if (float_equal(run, 0.0) && float_equal(rise, 0.0)) continue;
      double const lhs = (px - ax) * rise;  // run_to_point * rise
      double const rhs = run * (py - ay);   // run * rise_to_point
      if (float_equal(rhs, lhs)) {
        double const minx = ax < bx ? ax : bx;
        double const maxx = ax < bx ? bx : ax;
        if (minx <= px && px <= maxx) on_edge = true;
      }
      bool const y1 = ay > py;
      if (y1 != y0 && ((lhs < rhs) != y1)) inside = !inside;
      y0 = y1;
      bx = ax;
      by = ay;
    }
  }
  return inside && !on_edge;
}

// The y of a ring's last vertex (clamped as above): the `b` end of its first
// effective edge, whose y starts the crossing chain.
__device__ __forceinline__ double ring_last_y(PipArgs const& a, uint32_t r)
{
  uint64_t ve = a.ring_offsets[r + 1];
  if (ve > a.num_vertices) ve = a.num_vertices;
  return a.poly_y[ve - 1];
}

// A polygon's ring range and its first ring's table entries, loaded ahead of its
// test: a polygon-major wave loads them when it starts its unit, so that their
// dependent loads overlap the point gathers instead of starting each test.
struct RingHint {
  uint32_t r0, r1, e0, ne;
  double amax, by0;
};

__device__ __forceinline__ RingHint ring_hint(PipArgs const& a, uint32_t poly)
{
  RingHint h{0u, 0u, 0u, 0u, 0.0, 0.0};
  h.r0 = uniform(a.poly_offsets[poly]);
  h.r1 = uniform(a.poly_offsets[poly + 1]);
  if (h.r1 > a.num_rings) h.r1 = static_cast<uint32_t>(a.num_rings);
  if (h.r0 < h.r1) {
    h.e0   = uniform(a.ring_offsets[h.r0]);
    h.ne   = uniform(a.ring_edges[h.r0]);
    h.amax = a.ring_amax[h.r0];
    if (h.ne > 0) h.by0 = ring_last_y(a, h.r0);
  }
  return h;
}

template <int U, int S = kStageEdges>
__device__ __forceinline__ void test_polygon(PipArgs const& a, uint32_t poly,
                                             double const (&px)[kSlots],
                                             double const (&py)[kSlots],
                                             uint64_t const (&valid_m)[kSlots], double pmax,
                                             uint64_t (&hit_m)[kSlots], bool& err,
                                             PipStage<S>* stage, uint32_t lg = 0,
                                             RingHint const* hint = nullptr)
{
  uint64_t inside_m[U], suspect_m[U];
#pragma unroll
  for (int u = 0; u < U; ++u) inside_m[u] = suspect_m[u] = 0;
  uint32_t r0 = 0, r1 = 0;
  if (poly >= a.num_polys) {
    err = true;
  } else if (hint) {
    r0 = hint->r0;
    r1 = hint->r1;
  } else {
    r0 = uniform(a.poly_offsets[poly]);
    r1 = uniform(a.poly_offsets[poly + 1]);
    if (r1 > a.num_rings) r1 = static_cast<uint32_t>(a.num_rings);
  }
  int const lane = lane_id();
  for (uint32_t r = r0; r < r1; ++r) {
    bool const hinted = hint && r == r0;
    uint32_t const e0 = hinted ? hint->e0 : uniform(a.ring_offsets[r]);
    uint32_t const ne = hinted ? hint->ne : uniform(a.ring_edges[r]);
    if (ne == 0) continue;
    LineEdge const* L = a.lines + e0;
    // one tolerance for the ring: s bounds |px| + |ax| and |py| + |ay| (the wave's
    // points, the ring's vertices); the line forms are scaled to n' <= 1
    double const s   = pmax + (hinted ? hint->amax : a.ring_amax[r]);
    double const tol = s < 0x1p400 ? fma(s, 0x1p-48, 0x1p-1000) : __builtin_inf();
    // y0 of the first edge is its `b` (the ring's last vertex); afterwards the
    // previous edge's `a` (is_point_in_polygon.cuh:339, :375-376).
    double const by0 = hinted ? hint->by0 : ring_last_y(a, r);
    uint64_t y0[U];
#pragma unroll
    for (int u = 0; u < U; ++u) y0[u] = ballot(by0 > py[u]);
    // The reference toggles on (y1 != y0) && ((lhs < rhs) != y1); the line form is
    // oriented so that t < 0 is (lhs < rhs) != y1 wherever the edge straddles.
    auto edge = [&](LineEdge const& ed) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        double const t    = fma(px[u], ed.rise, fma(-py[u], ed.run, ed.c));
        uint64_t const y1 = ballot(ed.ay > py[u]);
        uint64_t const cr = ballot(t < 0.0);
        suspect_m[u] |= ballot(!(fabs(t) > tol));
        inside_m[u] ^= (y1 ^ y0[u]) & cr;
        y0[u] = y1;
      }
    };
    if (U == 1 && lg != 0) {
      uint32_t const G  = 1u << lg;
      uint32_t const Q  = static_cast<uint32_t>(kWave) >> lg;  // lanes per copy
      uint32_t const gq = static_cast<uint32_t>(lane) >> (6 - lg);
      // y1 of the previous step; before edge 0 every copy holds the ring's b vertex
      uint64_t y1p = ballot(by0 > py[0]);
      // one step: copy g takes edge e + g; the vertex before it is the previous
      // copy's edge (Q lanes down), or for copy 0 the last copy's previous step
      auto step = [&](LineEdge const& ed) {
        double const t    = fma(px[0], ed.rise, fma(-py[0], ed.run, ed.c));
        uint64_t const y1 = ballot(ed.ay > py[0]);
        uint64_t const cr = ballot(t < 0.0);
        suspect_m[0] |= ballot(!(fabs(t) > tol));
        uint64_t const y0 = (y1 << Q) | (y1p >> (kWave - Q));
        inside_m[0] ^= (y1 ^ y0) & cr;
        y1p = y1;
      };
      for (uint32_t base = 0; base < ne; base += S) {
        uint32_t const n   = ne - base < S ? ne - base : S;
        uint32_t const np  = (n + G - 1) & ~(G - 1);  // <= S: S is a multiple of G
        double const ylast = L[base + n - 1].ay;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int i = 0; i < S; i += kWave)
          if (static_cast<uint32_t>(i + lane) < n) stage->e[i + lane] = L[base + i + lane];
        if (static_cast<uint32_t>(lane) < np - n)
          stage->e[n + lane] = LineEdge{ylast, 0.0, 0.0, __builtin_inf()};
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        LineEdge const* E = &stage->e[gq];
        uint32_t e        = 0;
        for (; e + 2 * G <= np; e += 2 * G) {
          step(E[e]);
          step(E[e + G]);
        }
        if (e < np) step(E[e]);
      }
      continue;
    }
    for (uint32_t base = 0; base < ne; base += S) {
      uint32_t const n = ne - base < S ? ne - base : S;
      // the previous batch has been read by every lane before it is overwritten,
      // and the new one is written before any lane reads it
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int i = 0; i < S; i += kWave)
        if (static_cast<uint32_t>(i + lane) < n) stage->e[i + lane] = L[base + i + lane];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      uint32_t e = 0;
      // unrolled so that the loop's scalar bookkeeping is amortised
      for (; e + 2 <= n; e += 2) {
        edge(stage->e[e]);
        edge(stage->e[e + 1]);
      }
      if (e < n) edge(stage->e[e]);
    }
  }
  if (U == 1 && lg != 0) {
    // the copies' parities and marks onto the first copy's lanes
    for (uint32_t sh = kWave / 2; sh >= (static_cast<uint32_t>(kWave) >> lg); sh >>= 1) {
      inside_m[0] ^= inside_m[0] >> sh;
      suspect_m[0] |= suspect_m[0] >> sh;
    }
  }
  // Suspect points: the reference's test over the whole polygon
  // (is_point_in_polygon.cuh:46-100 on the effective edges).
#pragma unroll
  for (int u = 0; u < U; ++u) {
    uint64_t const sus = suspect_m[u] & valid_m[u];
    uint64_t exact_m   = 0;
    if (sus) {
      bool const suspect = (sus >> lane) & 1ull;
      bool const in      = suspect && exact_in_polygon(a, r0, r1, px[u], py[u]);
      exact_m            = ballot(in);
    }
    hit_m[u] = ((inside_m[u] & ~sus) | exact_m) & valid_m[u];
  }
}

// Point filter. A point strictly outside its polygon's vertex box (PolyBox) is
// outside the polygon in the reference's arithmetic too, so it needs no test:
//  * py < miny or py > maxy: every vertex is on one side of the point's y, so no
//    edge straddles it (y1 == y0 for every edge) and nothing toggles.
//  * px < minx - d or px > maxx + d, d = 2^-48 (|px| + |py| + B) + 2^-300, B the
//    box's largest |coordinate| (at most 2^400 and every non-zero rise at least
//    2^-600 in magnitude, else the box is infinite and filters nothing), and
//    |px| + |py| + B < 2^402: on a straddling edge (a, b) the reference's
//    lhs = fl(fl(px - ax) * rise) and rhs = fl(run * fl(py - ay)) have fixed
//    signs or, where the signs agree, |fl(px - ax)| >= |run| (1 + 3u) and
//    |fl(py - ay)| <= |rise| by monotone rounding (the gap d exceeds 12 u |run|),
//    so |lhs| > |rhs| with lhs a normal number (|px - ax| >= 2^-301,
//    |rise| >= 2^-600). Hence left of the box every straddling edge toggles
//    (an even number in a closed ring), right of it none does; the on-edge test
//    needs px within an edge's x range, so it never fires either.
//  NaN coordinates fail every comparison, so such points are always tested.
__device__ __forceinline__ bool box_excludes(PolyBox const& b, double px, double py)
{
  double const bm = fmax(fmax(fabs(b.minx), fabs(b.maxx)), fmax(fabs(b.miny), fabs(b.maxy)));
  double const s3 = fabs(px) + fabs(py) + bm;
  double const d  = fma(s3, 0x1p-48, 0x1p-300);
  bool const out_x = s3 < 0x1p402 && (px < b.minx - d || px > b.maxx + d);
  return py < b.miny || py > b.maxy || out_x;
}

// Moves the kept points (keep_m, c of them) to slots 0 .. ceil(c / 64) - 1 in
// keep order, through the wave's LDS staging area; test_m becomes their masks.
__device__ __forceinline__ void pack_points(uint64_t const (&keep_m)[kSlots],
                                            double (&px)[kSlots], double (&py)[kSlots],
                                            uint32_t c, uint64_t (&test_m)[kSlots], double2* buf)
{
  int const lane = lane_id();
  // the previous polygon's edges have been read by every lane
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t pre = 0;
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    if ((keep_m[u] >> lane) & 1ull)
      buf[pre + __popcll(keep_m[u] & lanemask_lt())] = make_double2(px[u], py[u]);
    pre += static_cast<uint32_t>(__popcll(keep_m[u]));
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    uint32_t const slot = static_cast<uint32_t>(u * kWave + lane);
    bool const in       = slot < c;
    if (in) {
      double2 const p = buf[slot];
      px[u]           = p.x;
      py[u]           = p.y;
    }
    test_m[u] = ballot(in);
  }
}

// Hits of packed slots back to the points' own slots: kept point j (in keep
// order) is bit j % 64 of packed slot j / 64.
__device__ __forceinline__ void unpack_hits(uint64_t const (&keep_m)[kSlots],
                                            uint64_t (&hit_m)[kSlots])
{
  int const lane = lane_id();
  uint64_t h_in[kSlots];
#pragma unroll
  for (int u = 0; u < kSlots; ++u) h_in[u] = hit_m[u];
  uint32_t pre = 0;
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    uint32_t const j = pre + static_cast<uint32_t>(__popcll(keep_m[u] & lanemask_lt()));
    uint64_t h       = h_in[0];
#pragma unroll
    for (int v = 1; v < kSlots; ++v)
      if (j >= static_cast<uint32_t>(v * kWave)) h = h_in[v];
    bool const hit = ((keep_m[u] >> lane) & 1ull) && ((h >> (j % kWave)) & 1ull);
    pre += static_cast<uint32_t>(__popcll(keep_m[u]));
    hit_m[u] = ballot(hit);
  }
}

// Pairs [base, base + waves of the grid) per launch: the host launches slices of
// kPipPairsPerLaunch pairs, so a grid never exceeds 2^28 work-items.
constexpr uint64_t kPipPairsPerLaunch = uint64_t{1} << 22;

// Eight waves per SIMD: the compiler then keeps the kernel within 96 scalar
// registers (spilling some of the per-polygon state to vector-register lanes,
// none inside the edge loops); the extra wave hides more of the point gathers'
// latency (batched join: 10.2 -> 10.0 ms per launch).
// Workgroups are dealt to the 8 XCDs round-robin (block b to XCD b % 8), and each
// XCD has its own L2. Adjacent pairs usually share a leaf, whose points a wave
// finds in L2 only if the wave that gathered them ran on the same XCD. So block b
// takes logical block (b / 8m) * 8m + (b % 8) * m + (b / 8) % m: each XCD works
// through runs of m consecutive logical blocks (m = pip_xcd_run); blocks past the
// last whole group of 8m keep their own index.
__device__ __forceinline__ uint64_t xcd_block(uint64_t b, uint64_t grid, uint32_t m)
{
  uint64_t const g = 8ull * m;
  if (m <= 1 || b >= grid / g * g) return b;
  return (b / g) * g + (b % 8) * m + (b / 8) % m;
}

template <int W, int S>  // W waves per workgroup, S edges per staged batch
__global__ void __launch_bounds__(W * kWave) __attribute__((amdgpu_waves_per_eu(8)))
  pip_test_kernel(PipArgs a, uint64_t base, uint32_t xcd_run, uint64_t* __restrict__ masks,
                  uint2* __restrict__ chunk_info, uint32_t* __restrict__ bad)
{
  __shared__ PipStage<S> s_stage[W];
  PipStage<S>* const stage = &s_stage[threadIdx.x / kWave];
  uint64_t const blk = xcd_block(blockIdx.x, gridDim.x, xcd_run);
  uint64_t const kk  = base + (blk * static_cast<uint64_t>(W) + threadIdx.x / kWave);
  uint64_t const K  = a.num_pairs_dev ? *a.num_pairs_dev : a.num_pairs;
  if (kk >= K) return;
  uint32_t const k0 = uniform(static_cast<uint32_t>(kk));
  uint32_t const quad = uniform(a.pair_quad[k0]);
  // A wave takes the pairs of one leaf from k0 up to the run's end or the next
  // multiple of pairs_per_wave, whichever comes first: a leaf that overlaps many
  // polygons is spread over several waves instead of being walked by one.
  uint32_t const G = a.pairs_per_wave;
  if (k0 > 0 && k0 % G != 0 && uniform(a.pair_quad[k0 - 1]) == quad) return;
  // Errors are recorded at the end: a global store before the edge loads would
  // keep them off the scalar unit (the compiler must assume it may alias).
  bool err = false, cap_err = false;
  if (quad >= a.num_nodes) {
    if (lane_id() == 0) atomicOr(bad, 1u);
    return;
  }
  uint32_t const len    = uniform(a.length[quad]);
  uint32_t const first  = uniform(a.offset[quad]);
  uint32_t const pbase  = a.point_base ? uniform(a.point_base[quad]) : 0u;
  uint32_t const chunks = (len + kWave - 1) / kWave;
  uint64_t const slot0  = uniform(a.chunk_start[k0]);
  int const lane        = lane_id();
  for (uint32_t j0 = 0; j0 < chunks; j0 += kSlots) {
    uint32_t const nu = chunks - j0 < kSlots ? chunks - j0 : kSlots;
    double px[kSlots], py[kSlots];
    uint64_t valid_m[kSlots];
#pragma unroll
    for (int u = 0; u < kSlots; ++u) {
      px[u] = py[u] = 0.0;
      valid_m[u]    = 0;
      if (static_cast<uint32_t>(u) >= nu) continue;  // wave-uniform
      uint32_t const local = (j0 + u) * kWave + lane;
      bool valid           = local < len;
      uint32_t const pos   = first + local;
      if (valid) {
        if (pos >= a.num_points) {
          err   = true;
          valid = false;
        } else {
          uint32_t const idx = a.point_indices[pos] + pbase;
          if (idx >= a.num_points) {
            err   = true;
            valid = false;
          } else if (a.xy) {
            if (a.xy_vec) {
              double2 const p = reinterpret_cast<double2 const*>(a.xy)[idx];
              px[u]           = p.x;
              py[u]           = p.y;
            } else {
              px[u] = a.xy[2 * static_cast<uint64_t>(idx)];
              py[u] = a.xy[2 * static_cast<uint64_t>(idx) + 1];
            }
          } else {
            px[u] = a.x[idx];
            py[u] = a.y[idx];
          }
        }
      }
      valid_m[u] = ballot(valid);
    }
    double const pmax = wave_max_abs(px, py, nu);
    for (uint32_t k = k0; k < K; ++k) {
      if (k > k0 && (k % G == 0 || uniform(a.pair_quad[k]) != quad)) break;
      uint32_t const poly = uniform(a.pair_poly[k]);
      uint64_t hit_m[kSlots];
      // the slots the test runs over, and their points
      uint32_t ut = nu;
      uint64_t test_m[kSlots], keep_m[kSlots];
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        hit_m[u]  = 0;
        test_m[u] = valid_m[u];
        keep_m[u] = 0;
      }
      bool packed = false;
      if (a.poly_box && poly < a.num_polys) {
        PolyBox const fb = a.poly_box[poly];
        uint32_t c       = 0;
#pragma unroll
        for (int u = 0; u < kSlots; ++u) {
          if (valid_m[u])  // wave-uniform
            keep_m[u] = ballot(((valid_m[u] >> lane) & 1ull) && !box_excludes(fb, px[u], py[u]));
          c += static_cast<uint32_t>(__popcll(keep_m[u]));
          test_m[u] = keep_m[u];
        }
        uint32_t const need = (c + kWave - 1) / kWave;
        // Packing overwrites the leaf's points, so only the wave's last pair packs.
        bool const last = k + 1 >= K || (k + 1) % G == 0 || uniform(a.pair_quad[k + 1]) != quad;
        if (c == 0) {
          ut = 0;
        } else if (need < nu && last && c * sizeof(double2) <= sizeof(PipStage<S>)) {
          pack_points(keep_m, px, py, c, test_m, reinterpret_cast<double2*>(stage));
          ut     = need;
          packed = true;
        }
      }
      static_assert(kSlots == 2, "the switch below covers 1 and 2 slots");
      if (ut == 1)
        test_polygon<1, S>(a, poly, px, py, test_m, pmax, hit_m, err, stage);
      else if (ut == 2)
        test_polygon<2, S>(a, poly, px, py, test_m, pmax, hit_m, err, stage);
      if (packed) unpack_hits(keep_m, hit_m);
      uint64_t const slot = slot0 + static_cast<uint64_t>(k - k0) * chunks + j0;
      if (slot + nu > a.chunk_cap) {
        cap_err = true;  // only with a caller-sized chunk capacity
      } else if (lane < static_cast<int>(nu)) {
        masks[slot + lane]      = lane == 0 ? hit_m[0] : hit_m[1];
        // what the write kernel needs: the polygon and the chunk's first point position
        chunk_info[slot + lane] = make_uint2(poly, first + (j0 + lane) * kWave);
      }
    }
  }
  if (err) atomicOr(bad, 1u);
  if (cap_err && lane == 0) atomicOr(bad, 2u);
}

// --- 3b. polygon-major test ---------------------------------------------------------
// The leaf-major test above runs the edge loop once per pair: with small leaves and
// a box filter that keeps few points per pair (the batched workload: about 2.6
// leaves per polygon, 17% of their points in the polygon's box), most lanes idle
// and each pair pays the wave's fixed costs. Here the pairs are grouped by polygon
// (a counting sort: pm_count_kernel, a scan, pm_scatter_kernel) and one wave takes
// up to pm_group pairs of one polygon: it gathers each pair's leaf chunk by chunk,
// keeps the points inside the polygon's box, and queues them in its lanes; every
// 64 queued points (and at the end) it runs the polygon's edges over them
// (test_polygon, one slot) and ORs each hit into its pair's chunk mask
// (zeroed beforehand). The wave also writes its pairs' chunk records. The results
// are the leaf-major kernel's: each (point, polygon) is decided the same way, and
// the order of the pairs inside a polygon (set by atomics) changes nothing.
__global__ void __launch_bounds__(256)
  pm_count_kernel(PipArgs a, uint32_t* __restrict__ count, uint32_t* __restrict__ bad)
{
  uint64_t const K = a.num_pairs_dev ? *a.num_pairs_dev : a.num_pairs;
  for (uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < K;
       k += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t const p = a.pair_poly[k], q = a.pair_quad[k];
    if (p >= a.num_polys || q >= a.num_nodes) {
      atomicOr(bad, 1u);
      continue;
    }
    atomicAdd(&count[p], 1u);
  }
}

struct PmScanOp {
  uint32_t const* counts;
  uint32_t* start;
  uint32_t* cursor;
  __device__ uint64_t count(uint64_t p) const { return counts[p]; }
  __device__ void emit(uint64_t p, uint64_t prefix) const
  {
    start[p]  = static_cast<uint32_t>(prefix);
    cursor[p] = static_cast<uint32_t>(prefix);
  }
};

// Units: polygon p's pairs [start, start + count) in runs of at most G.
struct PmUnitOp {
  uint32_t const* counts;
  uint32_t const* start;
  uint4* units;
  uint32_t G;
  __device__ uint64_t count(uint64_t p) const { return (counts[p] + G - 1) / G; }
  // one record per unit, so that a wave finds its polygon and pairs in one load
  __device__ void emit(uint64_t p, uint64_t prefix) const
  {
    uint32_t const c = counts[p], n = (c + G - 1) / G;
    for (uint32_t t = 0; t < n; ++t)
      units[prefix + t] = make_uint4(start[p] + t * G, static_cast<uint32_t>(p),
                                     c - t * G < G ? c - t * G : G, 0u);
  }
};

// A pair in the polygon-major order, as its wave reads it in one load: the leaf's
// point count and first sorted position, the node's job base (0 without one) and
// the pair's first chunk slot; the count is kPmOver for a pair whose chunks pass
// the chunk capacity (not tested; the caller sees bit 2). Slots fit 32 bits: the
// order is used only with a chunk capacity below 2^32.
constexpr uint32_t kPmOver = 0xFFFFFFFFu;

__global__ void __launch_bounds__(256)
  pm_scatter_kernel(PipArgs a, uint32_t* __restrict__ cursor, uint4* __restrict__ order)
{
  uint64_t const K = a.num_pairs_dev ? *a.num_pairs_dev : a.num_pairs;
  for (uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < K;
       k += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t const p = a.pair_poly[k], q = a.pair_quad[k];
    if (p >= a.num_polys || q >= a.num_nodes) continue;
    uint32_t const len   = a.length[q];
    uint32_t const first = a.offset[q];
    uint32_t const base  = a.point_base ? a.point_base[q] : 0u;
    uint64_t const sl    = a.chunk_start[k];
    bool const over      = sl + (len + kWave - 1) / kWave > a.chunk_cap;
    order[atomicAdd(&cursor[p], 1u)] =
      make_uint4(over ? kPmOver : len, first, base, static_cast<uint32_t>(sl));
  }
}

constexpr uint32_t kNoPoint = 0xFFFFFFFFu;  // a lane with no point to fetch

// A wave's queue transfer slot: kept points land here in queue order.
struct PmQueue {
  double2 p[kWave];
  uint64_t tag[kWave];  // chunk slot << 6 | lane of the point in its chunk
};

// One unit: the cnt pairs of polygon `poly` at order positions [i, i + cnt).
template <int S>
__device__ __forceinline__ void pm_unit(PipArgs const& a, uint32_t i, uint32_t poly, uint32_t cnt,
                                        uint64_t* __restrict__ masks,
                                        uint2* __restrict__ chunk_info, PipStage<S>* stage,
                                        PmQueue* tq, bool& err, bool& cap_err)
{
  uint32_t const end = i + cnt;
  RingHint const rh  = ring_hint(a, poly);
  int const lane     = lane_id();
  bool const filter = a.poly_box != nullptr;
  PolyBox fb{};
  if (filter) fb = a.poly_box[poly];
  // the queue: lanes [0, qn) hold queued points
  double qx[kSlots], qy[kSlots];
  uint64_t qtag = 0;
  uint32_t qn   = 0;
  double amax_p = 0.0;  // largest |coordinate| of the points queued since the last flush
#pragma unroll
  for (int u = 0; u < kSlots; ++u) qx[u] = qy[u] = 0.0;
  auto flush = [&]() {
    uint64_t test_m[kSlots], hit_m[kSlots];
#pragma unroll
    for (int u = 0; u < kSlots; ++u) test_m[u] = hit_m[u] = 0;
    test_m[0] = qn >= kWave ? ~0ull : ((1ull << qn) - 1ull);
    double m = amax_p;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
    // at most half a wave of points: copies of them in lane groups
    uint32_t lg = 0;
    if (qn <= kWave / 2) {
      uint32_t const q = qn > 1 ? 32u - static_cast<uint32_t>(__builtin_clz(qn - 1)) : 0u;
      lg               = 6 - q;  // 2^q >= qn lanes per copy
      int const src    = lane & ((1 << q) - 1);
      qx[0]            = __shfl(qx[0], src, kWave);
      qy[0]            = __shfl(qy[0], src, kWave);
    }
    test_polygon<1, S>(a, poly, qx, qy, test_m, m, hit_m, err, stage, lg, &rh);
    if ((hit_m[0] >> lane) & 1ull)
      atomicOr(reinterpret_cast<unsigned long long*>(&masks[qtag >> 6]), 1ull << (qtag & 63u));
    qn     = 0;
    amax_p = 0.0;
  };
  // the unit's pairs, one per lane (their loads in flight together): leaf range,
  // job base and first chunk slot
  uint32_t const m = end - static_cast<uint32_t>(i);
  uint32_t u_len = 0, u_first = 0, u_base = 0, u_slot = 0;
  auto rl = [](uint32_t v, uint32_t j) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(j)));
  };
  for (uint32_t j0 = 0; j0 < m; j0 += kWave) {
    uint32_t const mj = m - j0 < kWave ? m - j0 : kWave;
    if (static_cast<uint32_t>(lane) < mj) {
      uint4 const r = a.pm_order[i + j0 + lane];
      u_len         = r.x;
      u_first       = r.y;
      u_base        = r.z;
      u_slot        = r.w;
      // a pair whose chunks pass the capacity is not tested (the caller sees bit 2)
      if (u_len == kPmOver) {
        cap_err = true;
        u_len   = 0;
      }
    }
    auto slot_of = [&](uint32_t j) { return static_cast<uint64_t>(rl(u_slot, j)); };
    // the pairs' chunk records
    for (uint32_t j = 0; j < mj; ++j) {
      uint32_t const len   = rl(u_len, j);
      uint32_t const first = rl(u_first, j);
      uint64_t const slot0 = slot_of(j);
      for (uint32_t c = lane; c * kWave < len; c += kWave)
        chunk_info[slot0 + c] = make_uint2(poly, first + c * kWave);
    }
    // The items (pair j, chunk c) in order, software-pipelined: while item A is
    // filtered and queued, the coordinates of item B and the point index of item
    // C are in flight.
    auto next = [&](uint32_t& j, uint32_t& c) {  // the item after (j, c); j == mj: none
      if (j >= mj) return;
      ++c;
      while (j < mj && c * kWave >= rl(u_len, j)) {
        ++j;
        c = 0;
      }
    };
    // Reads of another lane's metadata (readlane) happen here, with every lane
    // active, before the per-lane branch: a value the compiler reloads from scratch
    // inside a branch is restored in the branch's active lanes only.
    auto fetch_index = [&](uint32_t j, uint32_t c, uint32_t& idx) {  // kNoPoint: none
      idx = kNoPoint;
      if (j >= mj) return;
      uint32_t const len   = rl(u_len, j);
      uint32_t const first = rl(u_first, j);
      uint32_t const pb    = rl(u_base, j);
      uint32_t const local = c * kWave + lane;
      if (local < len) {
        uint32_t const pos_p = first + local;
        if (pos_p >= a.num_points) {
          err = true;
        } else {
          uint32_t const v = a.point_indices[pos_p] + pb;
          if (v >= a.num_points)
            err = true;
          else
            idx = v;
        }
      }
    };
    auto fetch_point = [&](uint32_t idx, double& px, double& py) {
      px = py = 0.0;
      if (idx == kNoPoint) return;
      if (a.xy) {
        if (a.xy_vec) {
          double2 const p = reinterpret_cast<double2 const*>(a.xy)[idx];
          px              = p.x;
          py              = p.y;
        } else {
          px = a.xy[2 * static_cast<uint64_t>(idx)];
          py = a.xy[2 * static_cast<uint64_t>(idx) + 1];
        }
      } else {
        px = a.x[idx];
        py = a.y[idx];
      }
    };
    // the first item: chunk 0 of the first pair with points
    uint32_t ja = 0, ca = 0;
    while (ja < mj && rl(u_len, ja) == 0) ++ja;
    uint32_t jb = ja, cb = ca;
    next(jb, cb);
    uint32_t idx_a, idx_b;
    fetch_index(ja, ca, idx_a);
    fetch_index(jb, cb, idx_b);
    double ax_, ay_;
    fetch_point(idx_a, ax_, ay_);
    while (ja < mj) {
      uint32_t jc = jb, cc = cb;
      next(jc, cc);
      uint32_t idx_c;
      fetch_index(jc, cc, idx_c);
      double bx_, by_;
      fetch_point(idx_b, bx_, by_);
      // item A
      bool keep = idx_a != kNoPoint;
      if (keep && filter) keep = !box_excludes(fb, ax_, ay_);
      uint64_t const km = ballot(keep);
      if (km != 0) {
        uint32_t const nk   = static_cast<uint32_t>(__popcll(km));
        uint32_t const rank = qn + static_cast<uint32_t>(__popcll(km & lanemask_lt()));
        uint64_t const tag  = ((slot_of(ja) + ca) << 6) | static_cast<uint64_t>(lane);
        // first part: the kept points that fit in the queue
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (keep && rank < kWave) {
          tq->p[rank]   = make_double2(ax_, ay_);
          tq->tag[rank] = tag;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t const fill = qn + nk < kWave ? qn + nk : kWave;
        if (static_cast<uint32_t>(lane) >= qn && static_cast<uint32_t>(lane) < fill) {
          double2 const q = tq->p[lane];
          qx[0]           = q.x;
          qy[0]           = q.y;
          qtag            = tq->tag[lane];
          amax_p          = fmax(amax_p, fmax(fabs(q.x), fabs(q.y)));
        }
        uint32_t const total = qn + nk;
        qn                   = fill;
        if (total >= kWave) {
          flush();
          // second part: the rest starts the next queue
          uint32_t const rest = total - kWave;
          if (rest > 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (keep && rank >= kWave) {
              tq->p[rank - kWave]   = make_double2(ax_, ay_);
              tq->tag[rank - kWave] = tag;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (static_cast<uint32_t>(lane) < rest) {
              double2 const q = tq->p[lane];
              qx[0]           = q.x;
              qy[0]           = q.y;
              qtag            = tq->tag[lane];
              amax_p          = fmax(amax_p, fmax(fabs(q.x), fabs(q.y)));
            }
            qn = rest;
          }
        }
      }
      // shift the pipeline
      ja    = jb;
      ca    = cb;
      jb    = jc;
      cb    = cc;
      idx_a = idx_b;
      idx_b = idx_c;
      ax_   = bx_;
      ay_   = by_;
    }
  }
  if (qn > 0) flush();
}

// One wave per unit of the unit list, in the XCD-run workgroup order (a
// polygon's units are adjacent in the list, so its edges are fetched by one XCD);
// the grid covers the most units the pairs can form, and waves past the count
// return at once. No occupancy bound: held to the 7 or 8 waves per SIMD its LDS
// would admit, the compiler spills vector registers to scratch; unbounded it
// keeps 5 or 6 waves without spills, which measured faster.
template <int W, int S>
__global__ void __launch_bounds__(W * kWave)
  pip_test_pm_kernel(PipArgs a, uint64_t base, uint32_t xcd_run, uint64_t* __restrict__ masks,
                     uint2* __restrict__ chunk_info, uint32_t* __restrict__ bad)
{
  __shared__ PipStage<S> s_stage[W];
  __shared__ PmQueue s_queue[W];
  PipStage<S>* const stage = &s_stage[threadIdx.x / kWave];
  PmQueue* const tq        = &s_queue[threadIdx.x / kWave];
  int const lane = lane_id();
  uint64_t const blk = xcd_block(blockIdx.x, gridDim.x, xcd_run);
  uint64_t const u   = base + blk * static_cast<uint64_t>(W) + threadIdx.x / kWave;
  if (u >= *a.pm_num_units) return;
  bool err = false, cap_err = false;
  uint4 const un = a.pm_units[u];
  if (uniform(un.y) >= a.num_polys) {
    atomicOr(bad, 1u);
    return;
  }
  pm_unit(a, uniform(un.x), uniform(un.y), uniform(un.z), masks, chunk_info, stage, tq, err,
          cap_err);
  if (err) atomicOr(bad, 1u);
  if (cap_err && lane == 0) atomicOr(bad, 2u);
}

// --- 4. scan of hit counts ----------------------------------------------------------
struct HitScanOp {
  uint64_t const* masks;
  uint64_t* hit_start;
  __device__ uint64_t count(uint64_t c) const { return static_cast<uint64_t>(__popcll(masks[c])); }
  __device__ void emit(uint64_t c, uint64_t prefix) const { hit_start[c] = prefix; }
};

// --- 5. write -------------------------------------------------------------------------
// A wave owns 64 consecutive chunks: lane i loads chunk c0 + i's mask, and for a
// non-empty one its first hit and info, in coalesced loads (one round trip for
// 64 chunks instead of one wave, and two dependent round trips, per chunk). The
// wave then writes the non-empty chunks one at a time, each chunk's values
// broadcast from its lane; a chunk's hits are its set lanes, in lane order.
// With num_chunks_dev the chunk count lives on the device, and hits at or past
// hit_cap are dropped and flagged (bit 2 of *bad) instead of written.
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int j)
{
  uint32_t const lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), j));
  uint32_t const hi =
    static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), j));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__global__ void __launch_bounds__(256)
  pip_write_kernel(uint64_t num_chunks, uint64_t const* __restrict__ masks,
                   uint2 const* __restrict__ chunk_info, uint64_t const* __restrict__ hit_start,
                   uint32_t* __restrict__ out_poly, uint32_t* __restrict__ out_point,
                   uint64_t const* num_chunks_dev, uint64_t hit_cap, uint32_t* bad)
{
  uint64_t const T     = num_chunks_dev ? *num_chunks_dev : num_chunks;
  uint64_t const waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  int const l          = lane_id();
  uint64_t const lt    = lanemask_lt();
  for (uint64_t c0 = (blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * kWave; c0 < T;
       c0 += waves * kWave) {
    uint64_t const c = c0 + static_cast<uint64_t>(l);
    uint64_t mask = 0, start = 0;
    uint2 info    = make_uint2(0u, 0u);
    if (c < T) mask = masks[c];
    if (mask != 0) {
      start = hit_start[c];
      info  = chunk_info[c];
    }
    uint64_t pending = ballot(mask != 0);
    while (pending) {  // wave-uniform
      int const j = __builtin_ctzll(pending);
      pending &= pending - 1;
      // every cross-lane read before the per-lane branch (all lanes active)
      uint64_t const m     = readlane64(mask, j);
      uint64_t const s0    = readlane64(start, j);
      uint32_t const poly  = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(info.x), j));
      uint32_t const first = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(info.y), j));
      if ((m >> l) & 1ull) {
        uint64_t const o = s0 + static_cast<uint64_t>(__popcll(m & lt));
        if (o < hit_cap) {
          out_poly[o]  = poly;
          out_point[o] = first + static_cast<uint32_t>(l);
        } else {
          atomicOr(bad, 2u);
        }
      }
    }
  }
}

// Launches the test over pairs [0, pairs) in slices, with the tuned workgroup
// size (2, 4 or 8 waves) and staging depth (64 or 128 edges): the leaf-major
// kernel, or with pip_polygon_group > 0 (and `pm` given) the polygon-major one.
template <int W, int S>
void launch_pip_test_ws(PipArgs const& a, uint64_t pairs, uint64_t* masks, uint2* chunk_info,
                        uint32_t* bad, bool pm, hipStream_t stream)
{
  uint32_t const xcd_run = static_cast<uint32_t>(tuning(Tuning::PipXcdRun));
  if (pm) {
    // a polygon with c pairs forms ceil(c / G) <= c / G + 1 units
    uint64_t const units = std::min<uint64_t>(pairs, (pairs + a.pm_group - 1) / a.pm_group +
                                                       a.num_polys);
    for (uint64_t b = 0; b < units; b += kPipPairsPerLaunch) {
      uint64_t const n = std::min<uint64_t>(units - b, kPipPairsPerLaunch);
      unsigned const grid = static_cast<unsigned>((n + W - 1) / W);
      pip_test_pm_kernel<W, S><<<grid, W * kWave, 0, stream>>>(a, b, xcd_run, masks, chunk_info,
                                                               bad);
      QPIP_LAUNCH_CHECK();
    }
    return;
  }
  for (uint64_t b = 0; b < pairs; b += kPipPairsPerLaunch) {
    uint64_t const n = std::min<uint64_t>(pairs - b, kPipPairsPerLaunch);
    unsigned const grid = static_cast<unsigned>((n + W - 1) / W);
    pip_test_kernel<W, S><<<grid, W * kWave, 0, stream>>>(a, b, xcd_run, masks, chunk_info, bad);
    QPIP_LAUNCH_CHECK();
  }
}

void launch_pip_test(PipArgs const& args, uint64_t pairs, uint64_t* masks, uint2* chunk_info,
                     uint32_t* bad, PmBuffers const* pmb, hipStream_t stream)
{
  PipArgs a = args;
  if (tuning(Tuning::PipBoxFilter) == 0) a.poly_box = nullptr;
  int64_t const group = tuning(Tuning::PipPolygonGroup);
  bool const pm = group > 0 && pmb != nullptr && a.num_polys > 0 &&
                  a.chunk_cap < (uint64_t{1} << 32);  // slots in 32 bits (PmPair)
  if (pm) {
    // polygon-major order of the valid pairs; masks start at zero (hits are ORed in)
    KernelTimer kt("pip_order", stream);
    fill_u32(pmb->count, a.num_polys, 0u, stream);
    fill_u32(reinterpret_cast<uint32_t*>(masks), 2 * a.chunk_cap, 0u, stream);
    pm_count_kernel<<<grid_for(pairs, 256, 4096), 256, 0, stream>>>(a, pmb->count, bad);
    QPIP_LAUNCH_CHECK();
    PmScanOp op{pmb->count, pmb->start, pmb->cursor};
    count_scan_emit<PmScanOp, uint64_t, uint64_t>(op, nullptr, a.num_polys, pmb->partials,
                                                  pmb->total, stream,
                                                  scan_blocks_for(a.num_polys), "pip_order_scan");
    PmUnitOp uop{pmb->count, pmb->start, pmb->units, static_cast<uint32_t>(group)};
    count_scan_emit<PmUnitOp, uint64_t, uint64_t>(uop, nullptr, a.num_polys, pmb->partials,
                                                  pmb->num_units, stream,
                                                  scan_blocks_for(a.num_polys), "pip_order_scan");
    pm_scatter_kernel<<<grid_for(pairs, 256, 4096), 256, 0, stream>>>(a, pmb->cursor, pmb->order);
    QPIP_LAUNCH_CHECK();
    a.pm_order     = pmb->order;
    a.pm_start     = pmb->start;
    a.pm_count     = pmb->count;
    a.pm_units     = pmb->units;
    a.pm_num_units = pmb->num_units;
    a.pm_group     = static_cast<uint32_t>(group);
  }
  KernelTimer kt("pip_test", stream);
  int64_t const w = tuning(Tuning::PipWavesPerBlock);
  bool const deep = tuning(Tuning::PipStageEdges) > 64;
#define QPIP_PIP_TEST(W)                                                                    \
  (deep ? launch_pip_test_ws<W, 2 * kWave>(a, pairs, masks, chunk_info, bad, pm, stream) \
        : launch_pip_test_ws<W, kWave>(a, pairs, masks, chunk_info, bad, pm, stream))
  if (w <= 2)
    QPIP_PIP_TEST(2);
  else if (w <= 4)
    QPIP_PIP_TEST(4);
  else
    QPIP_PIP_TEST(8);
#undef QPIP_PIP_TEST
}

// *out = min(*in, cap); bit 2 of *flags when *in exceeds cap.
__global__ void clamp_count_kernel(uint64_t const* in, uint64_t cap, uint64_t* out,
                                   uint32_t* flags)
{
  if (threadIdx.x != 0) return;
  uint64_t const v = *in;
  if (v > cap) atomicOr(flags, 2u);
  *out = v < cap ? v : cap;
}

// --- brute force: every point against every polygon ------------------------------
// point_in_polygon (detail/point_in_polygon.cuh:42-102): bit j of the result of point
// i is point i inside polygon j, for at most 31 polygons. One wave per 128 points
// (2 per lane); each polygon's edges are wave-uniform, as in the quadtree test.
__device__ __forceinline__ void load_point(PipArgs const& a, uint64_t i, double& px, double& py)
{
  if (a.xy) {
    if (a.xy_vec) {
      double2 const p = reinterpret_cast<double2 const*>(a.xy)[i];
      px              = p.x;
      py              = p.y;
    } else {
      px = a.xy[2 * i];
      py = a.xy[2 * i + 1];
    }
  } else {
    px = a.x[i];
    py = a.y[i];
  }
}

__global__ void __launch_bounds__(256)
  pip_bitmask_kernel(PipArgs a, uint64_t n, int32_t* __restrict__ out)
{
  __shared__ PipStage<kStageEdges> s_stage[256 / kWave];
  PipStage<kStageEdges>* const stage = &s_stage[threadIdx.x / kWave];
  uint64_t const w  = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  uint64_t const p0 = w * kSlots * kWave;
  if (p0 >= n) return;
  int const lane    = lane_id();
  uint64_t const left = (n - p0 + kWave - 1) / kWave;
  uint32_t const nu   = left < kSlots ? static_cast<uint32_t>(left) : kSlots;
  double px[kSlots], py[kSlots];
  uint64_t valid_m[kSlots];
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    uint64_t const i = p0 + static_cast<uint64_t>(u) * kWave + lane;
    bool const valid = i < n;
    px[u] = py[u] = 0.0;
    if (valid) load_point(a, i, px[u], py[u]);
    valid_m[u] = ballot(valid);
  }
  double const pmax = wave_max_abs(px, py, nu);
  uint32_t bits[kSlots];
#pragma unroll
  for (int u = 0; u < kSlots; ++u) bits[u] = 0;
  bool err = false;
  for (uint32_t j = 0; j < a.num_polys; ++j) {
    uint64_t hit_m[kSlots];
#pragma unroll
    for (int u = 0; u < kSlots; ++u) hit_m[u] = 0;
    if (nu == 1)
      test_polygon<1>(a, j, px, py, valid_m, pmax, hit_m, err, stage);
    else
      test_polygon<2>(a, j, px, py, valid_m, pmax, hit_m, err, stage);
#pragma unroll
    for (int u = 0; u < kSlots; ++u) bits[u] |= static_cast<uint32_t>((hit_m[u] >> lane) & 1ull) << j;
  }
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    uint64_t const i = p0 + static_cast<uint64_t>(u) * kWave + lane;
    if (i < n) out[i] = static_cast<int32_t>(bits[u]);
  }
}

// pairwise_point_in_polygon (detail/point_in_polygon.cuh:104-160): point i against
// polygon i. Each lane has its own polygon, so the edges are read per lane.
__device__ __forceinline__ void pairwise_one(PipArgs const& a, uint64_t i, uint8_t* out)
{
  double px, py;
  load_point(a, i, px, py);
  uint32_t r0 = a.poly_offsets[i], r1 = a.poly_offsets[i + 1];
  if (r1 > a.num_rings) r1 = static_cast<uint32_t>(a.num_rings);
  out[i] = exact_in_polygon(a, r0, r1, px, py) ? 1 : 0;
}

__global__ void __launch_bounds__(256)
  pip_pairwise_kernel(PipArgs a, uint64_t n, uint8_t* __restrict__ out)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    pairwise_one(a, i, out);
}

}  // namespace

uint32_t pip_pairs_per_wave(uint64_t num_pairs) { return pairs_per_wave(num_pairs); }

void launch_pip_device(PipDevice const& d, hipStream_t stream)
{
  uint64_t* const n_pairs  = &d.counts->pairs;
  uint64_t* const n_chunks = &d.counts->chunks;
  uint32_t* const flags    = &d.counts->flags;
  ChunkOp cop{d.pair_quad, d.length, d.num_nodes, d.chunk_start, flags};
  count_scan_emit<ChunkOp, uint64_t, uint64_t>(cop, n_pairs, 0, d.partials, &d.totals[0], stream,
                                               kScanBlocks, "pip_chunks");
  clamp_count_kernel<<<1, 64, 0, stream>>>(&d.totals[0], d.chunk_cap, n_chunks, flags);
  QPIP_LAUNCH_CHECK();
  PipArgs args{d.pair_poly, d.pair_quad, d.pair_cap, d.length, d.offset, d.point_indices,
               d.num_points, nullptr, nullptr, d.poly_offsets, d.num_polys, d.num_rings,
               d.ring_offsets, d.ring_edges, d.lines, d.ring_amax,
               d.chunk_start, d.chunk_cap,
               d.num_nodes, d.xy, reinterpret_cast<uintptr_t>(d.xy) % 16 == 0, d.pairs_per_wave,
               n_pairs, d.chunk_cap, d.point_base, d.poly_box};
  args.poly_x       = d.poly_x;
  args.poly_y       = d.poly_y;
  args.num_vertices = d.num_vertices;
  QPIP_EXPECTS(args.num_vertices == 0 || (args.poly_x && args.poly_y),
               "polygon vertices must not be null");
  PmBuffers const pmb{d.pm_count,    d.pm_start, d.pm_cursor, d.pm_order,
                      d.pm_units,    d.pm_partials, d.pm_total, d.pm_num_units};
  launch_pip_test(args, d.pair_cap, d.masks, d.chunk_info, flags, d.pm_order ? &pmb : nullptr,
                  stream);
  HitScanOp hop{d.masks, d.hit_start};
  count_scan_emit<HitScanOp, uint64_t, uint64_t>(hop, n_chunks, 0, d.partials, &d.totals[1],
                                                 stream, kScanBlocks, "pip_scan");
  clamp_count_kernel<<<1, 64, 0, stream>>>(&d.totals[1], d.hit_cap, &d.counts->hits, flags);
  QPIP_LAUNCH_CHECK();
  {
    KernelTimer kt("pip_write", stream);
    // 4 waves per block, 64 chunks per wave; the kernel strides over what a
    // grid of at most 2048 blocks does not cover (a capacity may be far above
    // the count)
    pip_write_kernel<<<grid_for(d.chunk_cap, 4 * kWave, 2048), 256, 0, stream>>>(
      0, d.masks, d.chunk_info, d.hit_start, d.out_poly, d.out_pos, n_chunks, d.hit_cap, flags);
    QPIP_LAUNCH_CHECK();
  }
}

// The plain test's entry, also used by the batched eager call (point_base non-null:
// point_indices are job-local and point_base[node] is the node's job's first point).
int quadtree_point_in_polygon_entry(const uint32_t* poly_index,
                                              const uint32_t* quad_index,
                                              int64_t num_pairs,
                                              const uint32_t* key,
                                              const uint8_t* level,
                                              const uint8_t* is_internal_node,
                                              const uint32_t* length,
                                              const uint32_t* offset,
                                              int64_t num_nodes,
                                              const uint32_t* point_indices,
                                              int64_t num_point_indices,
                                              const double* x,
                                              const double* y,
                                              const double* xy,
                                              int64_t num_x,
                                              int64_t num_y,
                                              const uint32_t* poly_offsets,
                                              int64_t num_poly_offsets,
                                              const uint32_t* ring_offsets,
                                              int64_t num_ring_offsets,
                                              const double* poly_x,
                                              const double* poly_y,
                                              int64_t num_poly_x,
                                              int64_t num_poly_y,
                                              const qpip_resource* mr,
                                              void* stream_,
                                              uint32_t** polygon_index,
                                              uint32_t** point_index,
                                              int64_t* num_hits,
                                              const uint32_t* point_base)
{
  (void)key;
  (void)level;
  (void)is_internal_node;
  try {
    QPIP_EXPECTS(polygon_index && point_index && num_hits, "output pointers must not be null");
    *polygon_index = *point_index = nullptr;
    *num_hits                     = 0;
    hipStream_t stream            = static_cast<hipStream_t>(stream_);
    // quadtree_point_in_polygon.cu:154-178 (dtype checks are the binding's)
    QPIP_EXPECTS(num_point_indices == num_x && num_x == num_y,
                 "number of points must be the same for both x and y columns");
    QPIP_EXPECTS(num_poly_x == num_poly_y,
                 "numbers of vertices must be the same for both x and y columns");
    if (num_pairs == 0 || num_nodes == 0 || num_point_indices == 0 || num_poly_offsets == 0)
      return QPIP_SUCCESS;
    QPIP_EXPECTS(num_pairs < (int64_t{1} << 32) && num_x < (int64_t{1} << 32),
                 "inputs must have fewer than 2^32 rows");
    uint64_t const K         = static_cast<uint64_t>(num_pairs);
    uint64_t const num_rings = num_ring_offsets > 0 ? static_cast<uint64_t>(num_ring_offsets - 1) : 0;
    StreamScope scope(stream);

    // bad[0]: a quadrant index past the tree (chunk count); bad[1]: bit 1 a point or
    // polygon index out of range (test), bit 2 more chunks than the capacity
    dbuf<uint32_t> bad(2);
    dbuf<uint64_t> totals(2);  // chunks, hits
    QPIP_HIP(hipMemsetAsync(bad.get(), 0, 2 * sizeof(uint32_t), stream));

    // 1. edges
    dbuf<uint32_t> ring_edges(std::max<uint64_t>(num_rings, 1));
    dbuf<LineEdge> lines(std::max<int64_t>(num_poly_x, 1));
    dbuf<double> ring_amax(std::max<uint64_t>(num_rings, 1));
    uint64_t const num_polys = static_cast<uint64_t>(num_poly_offsets - 1);
    dbuf<PolyBox> ring_box(std::max<uint64_t>(num_rings, 1));
    dbuf<PolyBox> poly_box(std::max<uint64_t>(num_polys, 1));
    {
      KernelTimer kt("pip_edges", stream);
      if (num_rings > 0) {
        build_edges_kernel<<<grid_for(num_rings, 256 / kWave), 256, 0, stream>>>(
          ring_offsets, static_cast<int64_t>(num_rings), poly_x, poly_y, num_poly_x, nullptr,
          nullptr, ring_edges.get(), lines.get(), ring_amax.get(), ring_box.get());
        QPIP_LAUNCH_CHECK();
      }
      if (num_polys > 0) {
        poly_box_kernel<<<grid_for(num_polys, 256), 256, 0, stream>>>(
          poly_offsets, num_polys, num_rings, ring_box.get(), poly_box.get());
        QPIP_LAUNCH_CHECK();
      }
    }
    // 2. chunks: pair k's first chunk
    dbuf<uint32_t> chunk_start(K + 1);
    dbuf<uint64_t> partials(kScanBlocks + 1);
    ChunkOp cop{quad_index, length, static_cast<uint64_t>(num_nodes), chunk_start.get(), bad.get()};
    count_scan_emit<ChunkOp, uint64_t, uint64_t>(cop, nullptr, K, partials.get(), totals.get(),
                                                 stream, scan_blocks_for(K), "pip_chunks");

    // 3-4. test and hit scan into per-chunk arrays of a guessed capacity, so that
    // one host wait reads the chunk and hit totals together. The guess covers four
    // candidate points per input point (the headline has 0.4); a join with more
    // candidates is run again with the exact chunk count, as is a guess the
    // memory limit refuses.
    uint64_t const n_pts = static_cast<uint64_t>(num_x);
    uint64_t chunk_cap   = K + (n_pts >> 4) + 4096;
    bool exact           = false;
    dbuf<uint64_t> masks, hit_start;
    dbuf<uint2> chunk_info;
    // the polygon-major test's order (pip_polygon_group > 0)
    bool const pm_on = tuning(Tuning::PipPolygonGroup) > 0 && num_polys > 0;
    dbuf<uint32_t> pm_count(pm_on ? num_polys : 1), pm_start(pm_on ? num_polys : 1),
      pm_cursor(pm_on ? num_polys : 1);
    dbuf<uint4> pm_order(pm_on ? K : 1), pm_units(pm_on ? K : 1);
    dbuf<uint64_t> pm_partials(kScanBlocks + 1), pm_total(1), pm_num_units(1);
    PmBuffers const pmb{pm_count.get(), pm_start.get(),    pm_cursor.get(), pm_order.get(),
                        pm_units.get(), pm_partials.get(), pm_total.get(),  pm_num_units.get()};
    uint64_t h_tot[2] = {0, 0};
    uint32_t h_bad[2] = {0, 0};
    for (;;) {
      try {
        masks.reset(chunk_cap);
        chunk_info.reset(chunk_cap);
        hit_start.reset(chunk_cap);
      } catch (oom_error const&) {
        if (exact) throw;
        masks.release();
        chunk_info.release();
        hit_start.release();
        QPIP_HIP(hipMemcpyAsync(h_tot, totals.get(), sizeof(uint64_t), hipMemcpyDeviceToHost,
                                stream));
        QPIP_HIP(hipStreamSynchronize(stream));
        QPIP_EXPECTS(h_tot[0] < (uint64_t{1} << 31), "too many candidate points");
        chunk_cap = std::max<uint64_t>(h_tot[0], 1);
        exact     = true;
        continue;
      }
      PipArgs args{poly_index, quad_index, K, length, offset, point_indices,
                   n_pts, x, y, poly_offsets, num_polys, num_rings, ring_offsets,
                   ring_edges.get(), lines.get(), ring_amax.get(),
                   chunk_start.get(), chunk_cap,
                   static_cast<uint64_t>(num_nodes), xy,
                   xy != nullptr && reinterpret_cast<uintptr_t>(xy) % 16 == 0,
                   pairs_per_wave(K), nullptr, chunk_cap, point_base, poly_box.get()};
      args.poly_x       = poly_x;
      args.poly_y       = poly_y;
      args.num_vertices = static_cast<uint64_t>(num_poly_x);
      launch_pip_test(args, K, masks.get(), chunk_info.get(), bad.get() + 1,
                      pm_on ? &pmb : nullptr, stream);
      // chunks past the capacity were not tested: the scan covers those that were
      dbuf<uint64_t> n_chunks(1);
      clamp_count_kernel<<<1, 64, 0, stream>>>(totals.get(), chunk_cap, n_chunks.get(),
                                               bad.get() + 1);
      QPIP_LAUNCH_CHECK();
      HitScanOp hop{masks.get(), hit_start.get()};
      count_scan_emit<HitScanOp, uint64_t, uint64_t>(hop, n_chunks.get(), 0, partials.get(),
                                                     totals.get() + 1, stream, kScanBlocks,
                                                     "pip_scan");
      QPIP_HIP(hipMemcpyAsync(h_tot, totals.get(), sizeof(h_tot), hipMemcpyDeviceToHost, stream));
      QPIP_HIP(hipMemcpyAsync(h_bad, bad.get(), sizeof(h_bad), hipMemcpyDeviceToHost, stream));
      QPIP_HIP(hipStreamSynchronize(stream));
      if (h_bad[0]) throw logic_error("cuSpatial failure: quadrant index out of range of the quadtree");
      if (h_bad[1] & 1u) throw logic_error("cuSpatial failure: point or polygon index out of range");
      if (!(h_bad[1] & 2u)) break;
      // an exact capacity cannot overflow; past 2^31 chunks the starts saturate
      QPIP_EXPECTS(!exact && h_tot[0] < (uint64_t{1} << 31), "too many candidate points");
      chunk_cap = std::max<uint64_t>(h_tot[0], 1);
      exact     = true;
      QPIP_HIP(hipMemsetAsync(bad.get() + 1, 0, sizeof(uint32_t), stream));
    }
    uint64_t const T = h_tot[0];
    if (T == 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(T < (uint64_t{1} << 31), "too many candidate points");
    uint64_t const H = h_tot[1];
    if (H == 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(H < (uint64_t{1} << 32), "more than 2^32 hits");
    // 5. write (no wait: the outputs are ready for work queued on `stream` after
    // this call)
    OutputAllocator oa(mr, stream);
    uint32_t* o_poly  = oa.make<uint32_t>(H);
    uint32_t* o_point = oa.make<uint32_t>(H);
    {
      KernelTimer kt("pip_write", stream);
      pip_write_kernel<<<grid_for(T, 4 * kWave), 256, 0, stream>>>(T, masks.get(), chunk_info.get(),
                                                   hit_start.get(), o_poly, o_point, nullptr, H,
                                                   bad.get() + 1);
      QPIP_LAUNCH_CHECK();
    }
    oa.keep();
    *polygon_index = o_poly;
    *point_index   = o_point;
    *num_hits      = static_cast<int64_t>(H);
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}

}  // namespace qpip

extern "C" int qpip_quadtree_point_in_polygon(const uint32_t* poly_index,
                                              const uint32_t* quad_index,
                                              int64_t num_pairs,
                                              const uint32_t* key,
                                              const uint8_t* level,
                                              const uint8_t* is_internal_node,
                                              const uint32_t* length,
                                              const uint32_t* offset,
                                              int64_t num_nodes,
                                              const uint32_t* point_indices,
                                              int64_t num_point_indices,
                                              const double* x,
                                              const double* y,
                                              int64_t num_x,
                                              int64_t num_y,
                                              const uint32_t* poly_offsets,
                                              int64_t num_poly_offsets,
                                              const uint32_t* ring_offsets,
                                              int64_t num_ring_offsets,
                                              const double* poly_x,
                                              const double* poly_y,
                                              int64_t num_poly_x,
                                              int64_t num_poly_y,
                                              const qpip_resource* mr,
                                              void* stream,
                                              uint32_t** polygon_index,
                                              uint32_t** point_index,
                                              int64_t* num_hits)
{
  return qpip::quadtree_point_in_polygon_entry(
    poly_index, quad_index, num_pairs, key, level, is_internal_node, length, offset, num_nodes,
    point_indices, num_point_indices, x, y, nullptr, num_x, num_y, poly_offsets,
    num_poly_offsets, ring_offsets, num_ring_offsets, poly_x, poly_y, num_poly_x, num_poly_y, mr,
    stream, polygon_index, point_index, num_hits, nullptr);
}

extern "C" int qpip_quadtree_point_in_polygon_xy(const uint32_t* poly_index,
                                                 const uint32_t* quad_index,
                                                 int64_t num_pairs,
                                                 const uint32_t* key,
                                                 const uint8_t* level,
                                                 const uint8_t* is_internal_node,
                                                 const uint32_t* length,
                                                 const uint32_t* offset,
                                                 int64_t num_nodes,
                                                 const uint32_t* point_indices,
                                                 int64_t num_point_indices,
                                                 const double* points_xy,
                                                 int64_t num_points,
                                                 const uint32_t* poly_offsets,
                                                 int64_t num_poly_offsets,
                                                 const uint32_t* ring_offsets,
                                                 int64_t num_ring_offsets,
                                                 const double* poly_x,
                                                 const double* poly_y,
                                                 int64_t num_poly_x,
                                                 int64_t num_poly_y,
                                                 const qpip_resource* mr,
                                                 void* stream,
                                                 uint32_t** polygon_index,
                                                 uint32_t** point_index,
                                                 int64_t* num_hits)
{
  return qpip::quadtree_point_in_polygon_entry(
    poly_index, quad_index, num_pairs, key, level, is_internal_node, length, offset, num_nodes,
    point_indices, num_point_indices, nullptr, nullptr, points_xy, num_points, num_points,
    poly_offsets, num_poly_offsets, ring_offsets, num_ring_offsets, poly_x, poly_y, num_poly_x,
    num_poly_y, mr, stream, polygon_index, point_index, num_hits, nullptr);
}

namespace {

// cpp/src/point_in_polygon/point_in_polygon.cu:105-148 (column API checks), shared by
// the bitmask and pairwise entry points.
int point_in_polygon_entry(const double* x,
                           const double* y,
                           const double* xy,
                           int64_t num_x,
                           int64_t num_y,
                           const int32_t* poly_offsets,
                           int64_t num_poly_offsets,
                           const int32_t* ring_offsets,
                           int64_t num_ring_offsets,
                           const double* poly_x,
                           const double* poly_y,
                           int64_t num_poly_x,
                           int64_t num_poly_y,
                           const qpip_resource* mr,
                           void* stream_,
                           bool pairwise,
                           void** result,
                           int64_t* num_results)
{
  using namespace qpip;
  try {
    QPIP_EXPECTS(result && num_results, "output pointers must not be null");
    *result            = nullptr;
    *num_results       = 0;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    QPIP_EXPECTS(num_x == num_y && num_poly_x == num_poly_y,
                 "All points must have both x and y values");
    int64_t const num_polys = num_poly_offsets > 0 ? num_poly_offsets - 1 : 0;
    if (pairwise) {
      QPIP_EXPECTS(num_x == num_polys, "Must pass in the same number of points as polygons.");
    } else {
      // detail/point_in_polygon.cuh:94-95
      QPIP_EXPECTS(num_polys <= 31, "Number of polygons cannot exceed 31");
    }
    if (num_x <= 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(num_x < (int64_t{1} << 32), "inputs must have fewer than 2^32 rows");
    uint64_t const n         = static_cast<uint64_t>(num_x);
    uint64_t const num_rings = num_ring_offsets > 0 ? static_cast<uint64_t>(num_ring_offsets - 1) : 0;
    if (num_polys > 0) QPIP_EXPECTS(poly_offsets && ring_offsets, "offsets must not be null");
    StreamScope scope(stream);

    dbuf<uint32_t> ring_edges(std::max<uint64_t>(num_rings, 1));
    dbuf<LineEdge> lines(pairwise ? 1 : std::max<int64_t>(num_poly_x, 1));
    dbuf<double> ring_amax(pairwise ? 1 : std::max<uint64_t>(num_rings, 1));
    if (num_rings > 0) {
      KernelTimer kt("pip_edges", stream);
      build_edges_kernel<<<grid_for(num_rings, 256 / kWave), 256, 0, stream>>>(
        reinterpret_cast<uint32_t const*>(ring_offsets), static_cast<int64_t>(num_rings), poly_x,
        poly_y, num_poly_x, nullptr, nullptr, ring_edges.get(),
        pairwise ? nullptr : lines.get(), pairwise ? nullptr : ring_amax.get(), nullptr);
      QPIP_LAUNCH_CHECK();
    }
    PipArgs a{};
    a.num_points   = n;
    a.x            = x;
    a.y            = y;
    a.xy           = xy;
    a.xy_vec       = xy != nullptr && reinterpret_cast<uintptr_t>(xy) % 16 == 0;
    a.poly_offsets = reinterpret_cast<uint32_t const*>(poly_offsets);
    a.num_polys    = static_cast<uint64_t>(num_polys);
    a.num_rings    = num_rings;
    a.ring_offsets = reinterpret_cast<uint32_t const*>(ring_offsets);
    a.ring_edges   = ring_edges.get();
    a.poly_x       = poly_x;
    a.poly_y       = poly_y;
    a.num_vertices = static_cast<uint64_t>(num_poly_x);
    a.lines        = lines.get();
    a.ring_amax    = ring_amax.get();
    OutputAllocator oa(mr, stream);
    if (pairwise) {
      uint8_t* out = oa.make<uint8_t>(n);
      KernelTimer kt("pip_pairwise", stream);
      pip_pairwise_kernel<<<grid_for(n, 256), 256, 0, stream>>>(a, n, out);
      QPIP_LAUNCH_CHECK();
      *result = out;
    } else {
      int32_t* out = oa.make<int32_t>(n);
      KernelTimer kt("pip_bitmask", stream);
      uint64_t const waves = (n + kSlots * kWave - 1) / (kSlots * kWave);
      pip_bitmask_kernel<<<static_cast<unsigned>((waves + 3) / 4), 256, 0, stream>>>(a, n, out);
      QPIP_LAUNCH_CHECK();
      *result = out;
    }
    oa.keep();
    *num_results = num_x;
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}

}  // namespace

extern "C" int qpip_point_in_polygon(const double* x,
                                     const double* y,
                                     int64_t num_x,
                                     int64_t num_y,
                                     const int32_t* poly_offsets,
                                     int64_t num_poly_offsets,
                                     const int32_t* ring_offsets,
                                     int64_t num_ring_offsets,
                                     const double* poly_x,
                                     const double* poly_y,
                                     int64_t num_poly_x,
                                     int64_t num_poly_y,
                                     const qpip_resource* mr,
                                     void* stream,
                                     int32_t** result,
                                     int64_t* num_results)
{
  return point_in_polygon_entry(x, y, nullptr, num_x, num_y, poly_offsets, num_poly_offsets,
                                ring_offsets, num_ring_offsets, poly_x, poly_y, num_poly_x,
                                num_poly_y, mr, stream, false, reinterpret_cast<void**>(result),
                                num_results);
}

extern "C" int qpip_pairwise_point_in_polygon(const double* x,
                                              const double* y,
                                              int64_t num_x,
                                              int64_t num_y,
                                              const int32_t* poly_offsets,
                                              int64_t num_poly_offsets,
                                              const int32_t* ring_offsets,
                                              int64_t num_ring_offsets,
                                              const double* poly_x,
                                              const double* poly_y,
                                              int64_t num_poly_x,
                                              int64_t num_poly_y,
                                              const qpip_resource* mr,
                                              void* stream,
                                              uint8_t** result,
                                              int64_t* num_results)
{
  return point_in_polygon_entry(x, y, nullptr, num_x, num_y, poly_offsets, num_poly_offsets,
                                ring_offsets, num_ring_offsets, poly_x, poly_y, num_poly_x,
                                num_poly_y, mr, stream, true, reinterpret_cast<void**>(result),
                                num_results);
}

// --- allocation-free bitmask point-in-polygon ---------------------------------------
// qpip_point_in_polygon_noalloc runs the kernels of qpip_point_in_polygon with the
// edge table in `scratch`, a fixed device array of this file, instead of pool
// memory: the edges' line forms (32 B per vertex) and each ring's largest
// |coordinate| and edge count (12 B per ring); the exact retest reads the caller's
// vertices. 8.75 MiB per device, held by every device the library runs on until
// the process ends.
namespace {

constexpr size_t kNoAllocLinesOff  = 0;
constexpr size_t kNoAllocAmaxOff   = kNoAllocLinesOff + sizeof(qpip::LineEdge) * QPIP_NOALLOC_MAX_VERTICES;
constexpr size_t kNoAllocRingsOff  = kNoAllocAmaxOff + sizeof(double) * QPIP_NOALLOC_MAX_RINGS;
constexpr size_t kNoAllocBytes     = kNoAllocRingsOff + sizeof(uint32_t) * QPIP_NOALLOC_MAX_RINGS;
static_assert(kNoAllocLinesOff % alignof(qpip::LineEdge) == 0 && kNoAllocAmaxOff % 8 == 0,
              "scratch layout");
static_assert(kNoAllocRingsOff % 4 == 0, "scratch carving keeps each table aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each device, looked up once per device (with that
// device current) on the first call there. The host-side name `scratch` is not a
// device address and is never passed to a kernel or a copy.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

unsigned char* scratch_on_current_device()
{
  int dev = 0;
  QPIP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw qpip::device_error("device ordinal out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    QPIP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw qpip::device_error("scratch symbol has no device address");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" int qpip_point_in_polygon_noalloc(const double* x,
                                             const double* y,
                                             int64_t num_x,
                                             int64_t num_y,
                                             const int32_t* poly_offsets,
                                             int64_t num_poly_offsets,
                                             const int32_t* ring_offsets,
                                             int64_t num_ring_offsets,
                                             const double* poly_x,
                                             const double* poly_y,
                                             int64_t num_poly_x,
                                             int64_t num_poly_y,
                                             void* stream_,
                                             int32_t* result)
{
  using namespace qpip;
  try {
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    // the checks of qpip_point_in_polygon, in its order
    QPIP_EXPECTS(num_x == num_y && num_poly_x == num_poly_y,
                 "All points must have both x and y values");
    int64_t const num_polys = num_poly_offsets > 0 ? num_poly_offsets - 1 : 0;
    QPIP_EXPECTS(num_polys <= 31, "Number of polygons cannot exceed 31");
    if (num_x <= 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(num_x < (int64_t{1} << 32), "inputs must have fewer than 2^32 rows");
    QPIP_EXPECTS(result != nullptr, "result must not be null");
    uint64_t const num_rings = num_ring_offsets > 0 ? static_cast<uint64_t>(num_ring_offsets - 1) : 0;
    if (num_polys > 0) QPIP_EXPECTS(poly_offsets && ring_offsets, "offsets must not be null");
    if (num_poly_x > QPIP_NOALLOC_MAX_VERTICES || num_rings > QPIP_NOALLOC_MAX_RINGS)
      throw capacity_error(
        "qpip_point_in_polygon_noalloc: at most " QPIP_STR(QPIP_NOALLOC_MAX_VERTICES)
        " polygon vertices and " QPIP_STR(QPIP_NOALLOC_MAX_RINGS)
        " rings fit its fixed scratch; use qpip_point_in_polygon");
    unsigned char* const base = scratch_on_current_device();
    uint32_t* const ring_edges = reinterpret_cast<uint32_t*>(base + kNoAllocRingsOff);
    LineEdge* const lines      = reinterpret_cast<LineEdge*>(base + kNoAllocLinesOff);
    double* const ring_amax    = reinterpret_cast<double*>(base + kNoAllocAmaxOff);
    // Every ring's entry and every edge a ring keeps are written below before the
    // test reads them; the test reads nothing else of the scratch, so nothing a
    // previous call left there needs clearing.
    if (num_rings > 0) {
      build_edges_kernel<<<grid_for(num_rings, 256 / kWave), 256, 0, stream>>>(
        reinterpret_cast<uint32_t const*>(ring_offsets), static_cast<int64_t>(num_rings), poly_x,
        poly_y, num_poly_x, nullptr, nullptr, ring_edges, lines, ring_amax, nullptr);
      QPIP_LAUNCH_CHECK();
    }
    uint64_t const n = static_cast<uint64_t>(num_x);
    PipArgs a{};
    a.num_points   = n;
    a.x            = x;
    a.y            = y;
    a.poly_offsets = reinterpret_cast<uint32_t const*>(poly_offsets);
    a.num_polys    = static_cast<uint64_t>(num_polys);
    a.num_rings    = num_rings;
    a.ring_offsets = reinterpret_cast<uint32_t const*>(ring_offsets);
    a.ring_edges   = ring_edges;
    a.poly_x       = poly_x;
    a.poly_y       = poly_y;
    a.num_vertices = static_cast<uint64_t>(num_poly_x);
    a.lines        = lines;
    a.ring_amax    = ring_amax;
    uint64_t const waves = (n + kSlots * kWave - 1) / (kSlots * kWave);
    pip_bitmask_kernel<<<static_cast<unsigned>((waves + 3) / 4), 256, 0, stream>>>(a, n, result);
    QPIP_LAUNCH_CHECK();
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}
