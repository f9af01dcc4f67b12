// quadtree_point_to_nearest_linestring and linestring_bounding_boxes.
//
// Nearest linestring (detail/join/quadtree_point_to_nearest_linestring.cuh:143-314,
// cpp/src/join/quadtree_point_to_nearest_linestring.cu:128-189): for every point of a
// leaf that the (linestring, quadrant) join paired with linestrings, the distance
// to each of those linestrings (detail/algorithm/point_linestring_distance.cuh:33-51)
// and the one kept by the reference's reduce_by_key operator: the smallest nonzero
// distance, ties to the lower linestring id; a zero distance yields to any other
// (:279-294). That operator is associative (zero distances are identities unless
// every distance is zero, when the last pair of the run wins), so a left fold over
// the run's pairs in pair order gives the reference's result.
//
// Output has one row per entry of point_indices (the reference sizes its outputs
// by the point count, :257-264): row p is (p, nearest linestring, distance) for a
// point in a paired leaf. The reference leaves the two index columns of the other
// rows uninitialized (distance 0); here they are (p, 0xffffffff, 0).
//
// One wave per pair: the waves of a run (pairs on one leaf, adjacent in the
// join's order, which sorts pairs by leaf offset) split the leaf's 64-point chunks
// between them (chunk c goes to the
// run's pair c mod R); each lane holds one point and folds over the run's
// linestrings, whose vertices are wave-uniform.
#include <algorithm>
#include <cmath>

#include "qpip_internal.hpp"

namespace qpip {

namespace {

struct NearestArgs {
  uint32_t const* pair_ls;
  uint32_t const* pair_quad;
  uint64_t num_pairs;
  uint32_t const* length;
  uint32_t const* offset;
  uint64_t num_nodes;
  uint32_t const* point_indices;
  uint64_t num_points;
  double const* x;
  double const* y;
  uint32_t const* ls_offsets;
  uint64_t num_ls;
  double const* lx;
  double const* ly;
  uint64_t num_vertices;
  uint32_t* out_point;
  uint32_t* out_ls;
  double* out_dist;
  uint32_t* bad;
};

template <class T>
__device__ __forceinline__ T uniform_load(T const* p)
{
  return __builtin_amdgcn_readfirstlane(*p);
}

__device__ __forceinline__ double uniform_d(double const* p)
{
  uint64_t const b = static_cast<uint64_t>(__double_as_longlong(*p));
  uint32_t const lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b));
  uint32_t const hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b >> 32));
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// dot(a, b) = a.x * b.x + a.y * b.y as fma(a.x, b.x, a.y * b.y): the form that
// reproduces the reference's expected distances bit for bit
// (quadtree_point_to_nearest_linestring_test_small.cu:173-192).
__device__ __forceinline__ double dot2(double ax, double ay, double bx, double by)
{
  return fma(ax, bx, ay * by);
}

// point_linestring_distance (detail/algorithm/point_linestring_distance.cuh:33-51):
// min over the segments (v[i], v[i+1]) of the squared distance, then the square
// root.
__device__ __forceinline__ double point_linestring_distance(NearestArgs const& a, uint32_t ls,
                                                            double px, double py)
{
  uint64_t v0 = uniform_load(a.ls_offsets + ls);
  uint64_t v1 = uniform_load(a.ls_offsets + ls + 1);
  if (v1 > a.num_vertices) v1 = a.num_vertices;
  double dsq = 1.7976931348623157e308;  // numeric_limits<double>::max()
  for (uint64_t i = v0; i + 1 < v1; ++i) {
    double const ax = uniform_d(a.lx + i), ay = uniform_d(a.ly + i);
    double const bx = uniform_d(a.lx + i + 1), by = uniform_d(a.ly + i + 1);
    double const v1x = px - ax, v1y = py - ay;
    double const v2x = px - bx, v2y = py - by;
    double const ex = bx - ax, ey = by - ay;
    double const d0 = dot2(v1x, v1y, v1x, v1y);
    double const d1 = dot2(v2x, v2y, v2x, v2y);
    double const d2 = dot2(ex, ey, ex, ey);
    double const d3 = dot2(v1x, v1y, ex, ey);
    double const r  = d3 * d3 / d2;
    double const d  = (d3 <= 0.0 || r >= d2) ? fmin(d0, d1) : d0 - r;
    dsq             = fmin(dsq, d);
  }
  return sqrt(dsq);
}

__global__ void __launch_bounds__(256) nearest_fill_kernel(NearestArgs a)
{
  uint64_t const p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= a.num_points) return;
  a.out_point[p] = static_cast<uint32_t>(p);
  a.out_ls[p]    = 0xffffffffu;
  a.out_dist[p]  = 0.0;
}

// Pairs [base, base + waves of the grid): the host launches slices of 2^22 pairs so
// that a grid stays within 2^28 work-items.
__global__ void __launch_bounds__(256) nearest_kernel(NearestArgs a, uint64_t base)
{
  uint64_t const k0 = base + __builtin_amdgcn_readfirstlane(
    static_cast<uint32_t>(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
  if (k0 >= a.num_pairs) return;
  uint32_t const quad = uniform_load(a.pair_quad + k0);
  if (quad >= a.num_nodes) {
    if (lane_id() == 0) atomicOr(a.bad, 1u);
    return;
  }
  uint32_t const len   = uniform_load(a.length + quad);
  uint32_t const first = uniform_load(a.offset + quad);
  // The run of pairs on this leaf, found as the reference finds it
  // (quadtree_point_to_nearest_linestring.cuh:44-61): lower and upper bound of the
  // leaf's offset among the pairs' quadrant offsets (the join orders pairs by them).
  auto off_at = [&](uint64_t k) -> uint32_t {
    uint32_t const q = uniform_load(a.pair_quad + k);
    return q < a.num_nodes ? uniform_load(a.offset + q) : 0xffffffffu;
  };
  uint64_t lo = 0, hi = k0;
  while (lo < hi) {
    uint64_t const mid = (lo + hi) / 2;
    if (off_at(mid) < first) lo = mid + 1;
    else hi = mid;
  }
  uint64_t const rs = lo;
  lo = k0 + 1;
  hi = a.num_pairs;
  while (lo < hi) {
    uint64_t const mid = (lo + hi) / 2;
    if (off_at(mid) <= first) lo = mid + 1;
    else hi = mid;
  }
  uint64_t const re     = lo;
  uint64_t const R      = re - rs;
  uint64_t const mine   = k0 - rs;
  uint64_t const chunks = (static_cast<uint64_t>(len) + kWave - 1) / kWave;
  int const lane        = lane_id();
  bool err              = false;
  for (uint64_t c = mine; c < chunks; c += R) {
    uint64_t const local = c * kWave + lane;
    if (local >= len) continue;
    uint64_t const pos = static_cast<uint64_t>(first) + local;
    if (pos >= a.num_points) {
      err = true;
      continue;
    }
    uint32_t const idx = a.point_indices[pos];
    if (idx >= a.num_points) {
      err = true;
      continue;
    }
    double const px = a.x[idx], py = a.y[idx];
    uint32_t best_ls = 0;
    double best_d    = 0.0;
    for (uint64_t k = rs; k < re; ++k) {
      uint32_t const ls = uniform_load(a.pair_ls + k);
      if (ls >= a.num_ls) {
        err = true;
        continue;
      }
      double const d = point_linestring_distance(a, ls, px, py);
      if (k == rs) {
        best_ls = ls;
        best_d  = d;
        continue;
      }
      // quadtree_point_to_nearest_linestring.cuh:279-294, lhs = best, rhs = this pair
      bool take_rhs;
      if (best_d == 0.0) take_rhs = true;
      else if (d == 0.0) take_rhs = false;
      else if (best_d == d) take_rhs = !(best_ls < ls);
      else take_rhs = !(best_d < d);
      if (take_rhs) {
        best_ls = ls;
        best_d  = d;
      }
    }
    a.out_point[pos] = static_cast<uint32_t>(pos);
    a.out_ls[pos]    = best_ls;
    a.out_dist[pos]  = best_d;
  }
  if (err) atomicOr(a.bad, 1u);
}

__global__ void iota_i32_kernel(int32_t* __restrict__ p, int64_t n)
{
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = static_cast<int32_t>(i);
}

}  // namespace

}  // namespace qpip

using namespace qpip;

extern "C" int qpip_quadtree_point_to_nearest_linestring(const uint32_t* linestring_index,
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
                                                         const uint32_t* linestring_offsets,
                                                         int64_t num_linestring_offsets,
                                                         const double* linestring_x,
                                                         const double* linestring_y,
                                                         int64_t num_linestring_x,
                                                         int64_t num_linestring_y,
                                                         const qpip_resource* mr,
                                                         void* stream_,
                                                         uint32_t** point_index,
                                                         uint32_t** nearest_linestring_index,
                                                         double** distance,
                                                         int64_t* num_rows)
{
  (void)key;
  (void)level;
  (void)is_internal_node;
  try {
    QPIP_EXPECTS(point_index && nearest_linestring_index && distance && num_rows,
                 "output pointers must not be null");
    *point_index = *nearest_linestring_index = nullptr;
    *distance                                = nullptr;
    *num_rows                                = 0;
    auto stream = static_cast<hipStream_t>(stream_);
    // quadtree_point_to_nearest_linestring.cu:153-166 (column counts and dtypes are the
    // binding's: 2 + 5 columns, uint32 indices, f64 coordinates)
    QPIP_EXPECTS(num_point_indices == num_x && num_x == num_y,
                 "number of points must be the same for both x and y columns");
    QPIP_EXPECTS(num_linestring_x == num_linestring_y,
                 "numbers of vertices must be the same for both x and y columns");
    // :168-176: empty inputs give empty columns
    if (num_pairs == 0 || num_nodes == 0 || num_point_indices == 0 ||
        num_linestring_offsets == 0)
      return QPIP_SUCCESS;
    int64_t const n = num_point_indices;
    StreamScope scope(stream);
    OutputAllocator oa(mr, stream);
    uint32_t* o_point = oa.make<uint32_t>(n);
    uint32_t* o_ls    = oa.make<uint32_t>(n);
    double* o_dist    = oa.make<double>(n);
    dbuf<uint32_t> bad(1);
    fill_u32(bad.get(), 1, 0u, stream);
    NearestArgs a{linestring_index, quad_index, static_cast<uint64_t>(num_pairs), length, offset,
                  static_cast<uint64_t>(num_nodes), point_indices, static_cast<uint64_t>(n), x, y,
                  linestring_offsets, static_cast<uint64_t>(num_linestring_offsets - 1),
                  linestring_x, linestring_y, static_cast<uint64_t>(num_linestring_x), o_point,
                  o_ls, o_dist, bad.get()};
    {
      KernelTimer kt("nearest_fill", stream);
      nearest_fill_kernel<<<grid_for(static_cast<uint64_t>(n), 256, 1u << 30), 256, 0, stream>>>(a);
      QPIP_LAUNCH_CHECK();
    }
    {
      KernelTimer kt("nearest_linestring", stream);
      uint64_t const K = static_cast<uint64_t>(num_pairs);
      for (uint64_t b = 0; b < K; b += uint64_t{1} << 22) {
        uint64_t const m = std::min<uint64_t>(K - b, uint64_t{1} << 22);
        nearest_kernel<<<static_cast<unsigned>((m + 3) / 4), 256, 0, stream>>>(a, b);
        QPIP_LAUNCH_CHECK();
      }
    }
    uint32_t h_bad = 0;
    QPIP_HIP(hipMemcpyAsync(&h_bad, bad.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    QPIP_HIP(hipStreamSynchronize(stream));
    if (h_bad)  // the outputs are released by `oa`
      throw logic_error(
        "cuSpatial failure: a quadrant, point or linestring index is out of range");
    oa.keep();
    *point_index              = o_point;
    *nearest_linestring_index = o_ls;
    *distance                 = o_dist;
    *num_rows                 = n;
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}

extern "C" int qpip_linestring_bounding_boxes(const int32_t* linestring_offsets,
                                              int64_t num_linestring_offsets,
                                              const double* x,
                                              const double* y,
                                              int64_t num_x,
                                              int64_t num_y,
                                              double expansion_radius,
                                              const qpip_resource* mr,
                                              void* stream_,
                                              double** minx,
                                              double** miny,
                                              double** maxx,
                                              double** maxy,
                                              int64_t* num_boxes)
{
  try {
    QPIP_EXPECTS(minx && miny && maxx && maxy && num_boxes, "output pointers must not be null");
    *minx = *miny = *maxx = *maxy = nullptr;
    *num_boxes                    = 0;
    auto stream                   = static_cast<hipStream_t>(stream_);
    // linestring_bounding_boxes.cu:133-141
    int64_t const num_ls = num_linestring_offsets > 0 ? num_linestring_offsets - 1 : 0;
    QPIP_EXPECTS(num_x == num_y, "x and y must be the same size");
    QPIP_EXPECTS(expansion_radius >= 0, "expansion radius must be greater or equal than 0");
    QPIP_EXPECTS(num_x >= 2 * num_ls, "all linestrings must have at least 2 vertices");
    if (num_ls == 0 || num_x == 0) return QPIP_SUCCESS;
    // a linestring is a polygon with one ring: the same per-geometry reduction
    // (detail/bounding_boxes.cuh:97-131 and :139-184 both reduce by geometry id)
    StreamScope scope(stream);
    dbuf<int32_t> parts(num_ls + 1);
    iota_i32_kernel<<<grid_for(static_cast<uint64_t>(num_ls + 1), 256, 1u << 30), 256, 0,
                      stream>>>(parts.get(), num_ls + 1);
    QPIP_LAUNCH_CHECK();
    int const rc = qpip_polygon_bounding_boxes(parts.get(), num_ls + 1, linestring_offsets,
                                               num_linestring_offsets, x, y, num_x, num_y,
                                               expansion_radius, mr, stream_, minx, miny, maxx,
                                               maxy, num_boxes);
    return rc;
  } catch (...) {
    return qpip::error_status();
  }
}
