/*
 * C++ interface of libcuspatial_amd: the reference's C++ column API for the
 * quadtree point-in-polygon path, over the C ABI in cuspatial_amd.h.
 *
 * Mirrors (same functions, argument order, output columns, exceptions):
 *   cuspatial::quadtree_on_points                cpp/include/cuspatial/point_quadtree.hpp:68-78
 *   cuspatial::polygon_bounding_boxes            cpp/include/cuspatial/bounding_boxes.hpp:79-85
 *   cuspatial::linestring_bounding_boxes         cpp/include/cuspatial/bounding_boxes.hpp:52-57
 *   cuspatial::join_quadtree_and_bounding_boxes  cpp/include/cuspatial/spatial_join.hpp:66-75
 *   cuspatial::quadtree_point_in_polygon         cpp/include/cuspatial/spatial_join.hpp:116-126
 *   cuspatial::quadtree_point_to_nearest_linestring
 *                                                cpp/include/cuspatial/spatial_join.hpp:166-175
 *
 * Inputs are device arrays with their lengths (the reference takes
 * cudf::column_view); outputs are move-only device columns that own their
 * memory (the reference returns std::unique_ptr<cudf::column> / cudf::table).
 * Every function ends with the reference's stream and memory resource
 * arguments (cpp/include/cuspatial/spatial_join.cuh:73-74,
 * point_quadtree.cuh:179-180): `stream` (a hipStream_t; the work is ordered on
 * it and the outputs are ready for work queued on it after the call) and `mr`
 * (a qpip_resource, the counterpart of rmm::device_async_resource_ref: every
 * output column is allocated from it and released through it; nullptr means
 * the library's pool).
 * Violated preconditions throw cuspatial_amd::logic_error with the reference's
 * reason; device failures throw cuspatial_amd::device_error (the reference's
 * cuda_error); exhausted memory throws std::bad_alloc.
 *
 * Header only; link with libcuspatial_amd.so.
 */
#ifndef CUSPATIAL_AMD_HPP
#define CUSPATIAL_AMD_HPP

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "cuspatial_amd.h"

namespace cuspatial_amd {

struct logic_error : std::logic_error {
  using std::logic_error::logic_error;
};
struct device_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

namespace detail {
inline void check(int rc)
{
  if (rc == QPIP_SUCCESS) return;
  std::string const msg = qpip_last_error_message();
  if (rc == QPIP_ERROR_LOGIC) throw logic_error(msg);
  if (rc == QPIP_ERROR_OOM) throw std::bad_alloc();
  throw device_error(msg);
}
}  // namespace detail

/* A device column that owns its memory: allocated from the caller's resource
 * `mr` (released through it, on the stream of the call that made the column),
 * or, with mr == nullptr, from the library's pool (released with qpip_free). */
template <class T>
class column {
 public:
  column() = default;
  column(T* data, int64_t size, qpip_resource const* mr = nullptr, void* stream = nullptr)
    : data_(data), size_(size), mr_(mr), stream_(stream)
  {
  }
  ~column() { release(); }
  column(column const&)            = delete;
  column& operator=(column const&) = delete;
  column(column&& o) noexcept : data_(o.data_), size_(o.size_), mr_(o.mr_), stream_(o.stream_)
  {
    o.data_ = nullptr;
    o.size_ = 0;
  }
  column& operator=(column&& o) noexcept
  {
    if (this != &o) {
      release();
      data_   = o.data_;
      size_   = o.size_;
      mr_     = o.mr_;
      stream_ = o.stream_;
      o.data_ = nullptr;
      o.size_ = 0;
    }
    return *this;
  }
  T* data() const { return data_; }
  int64_t size() const { return size_; }
  qpip_resource const* resource() const { return mr_; }

 private:
  void release()
  {
    if (!data_) return;
    if (mr_)
      mr_->deallocate(mr_->ctx, data_, static_cast<size_t>(size_) * sizeof(T), stream_);
    else
      qpip_free(data_);
    data_ = nullptr;
  }
  T* data_                  = nullptr;
  int64_t size_             = 0;
  qpip_resource const* mr_  = nullptr;
  void* stream_             = nullptr;
};

/* The five quadtree columns (key UINT32, level UINT8, is_internal_node BOOL8,
 * length UINT32, offset UINT32). */
struct point_quadtree {
  column<uint32_t> key;
  column<uint8_t> level;
  column<uint8_t> is_internal_node;
  column<uint32_t> length;
  column<uint32_t> offset;
  int64_t size() const { return key.size(); }
};

struct bounding_boxes {
  column<double> minx, miny, maxx, maxy;
  int64_t size() const { return minx.size(); }
};

/* (bbox_offset, quad_offset) rows of join_quadtree_and_bounding_boxes. */
struct bbox_quad_pairs {
  column<uint32_t> bbox_offset, quad_offset;
  int64_t size() const { return bbox_offset.size(); }
};

/* (polygon_index, point_index) rows of quadtree_point_in_polygon. */
struct polygon_point_pairs {
  column<uint32_t> polygon_index, point_index;
  int64_t size() const { return polygon_index.size(); }
};

/* (point_index, linestring_index, distance) rows of
 * quadtree_point_to_nearest_linestring. */
struct point_linestring_distances {
  column<uint32_t> point_index, linestring_index;
  column<double> distance;
  int64_t size() const { return point_index.size(); }
};

/* quadtree_on_points(x, y, x_min, x_max, y_min, y_max, scale, max_depth,
 * max_size) -> (point_indices, quadtree) */
inline std::pair<column<uint32_t>, point_quadtree> quadtree_on_points(
  double const* x, int64_t num_x, double const* y, int64_t num_y, double x_min, double x_max,
  double y_min, double y_max, double scale, int8_t max_depth, int32_t max_size,
  void* stream = nullptr, qpip_resource const* mr = nullptr)
{
  uint32_t *pi = nullptr, *key = nullptr, *length = nullptr, *offset = nullptr;
  uint8_t *level = nullptr, *internal = nullptr;
  int64_t m = 0;
  detail::check(qpip_quadtree_on_points(x, y, num_x, num_y, x_min, x_max, y_min, y_max, scale,
                                        max_depth, max_size, mr, stream, &pi, &key, &level,
                                        &internal, &length, &offset, &m));
  point_quadtree t{column<uint32_t>(key, m, mr, stream), column<uint8_t>(level, m, mr, stream),
                   column<uint8_t>(internal, m, mr, stream),
                   column<uint32_t>(length, m, mr, stream),
                   column<uint32_t>(offset, m, mr, stream)};
  return {column<uint32_t>(pi, m > 0 ? num_x : 0, mr, stream), std::move(t)};
}

/* polygon_bounding_boxes(poly_offsets, ring_offsets, x, y, expansion_radius) */
inline bounding_boxes polygon_bounding_boxes(int32_t const* poly_offsets, int64_t num_poly_offsets,
                                             int32_t const* ring_offsets, int64_t num_ring_offsets,
                                             double const* x, int64_t num_x, double const* y,
                                             int64_t num_y, double expansion_radius = 0.0,
                                             void* stream = nullptr,
                                             qpip_resource const* mr = nullptr)
{
  double *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  int64_t n = 0;
  detail::check(qpip_polygon_bounding_boxes(poly_offsets, num_poly_offsets, ring_offsets,
                                            num_ring_offsets, x, y, num_x, num_y, expansion_radius,
                                            mr, stream, &a, &b, &c, &d, &n));
  return bounding_boxes{column<double>(a, n, mr, stream), column<double>(b, n, mr, stream),
                        column<double>(c, n, mr, stream), column<double>(d, n, mr, stream)};
}

/* linestring_bounding_boxes(linestring_offsets, x, y, expansion_radius) */
inline bounding_boxes linestring_bounding_boxes(int32_t const* linestring_offsets,
                                                int64_t num_linestring_offsets, double const* x,
                                                int64_t num_x, double const* y, int64_t num_y,
                                                double expansion_radius, void* stream = nullptr,
                                                qpip_resource const* mr = nullptr)
{
  double *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  int64_t n = 0;
  detail::check(qpip_linestring_bounding_boxes(linestring_offsets, num_linestring_offsets, x, y,
                                               num_x, num_y, expansion_radius, mr, stream, &a, &b,
                                               &c, &d, &n));
  return bounding_boxes{column<double>(a, n, mr, stream), column<double>(b, n, mr, stream),
                        column<double>(c, n, mr, stream), column<double>(d, n, mr, stream)};
}

/* join_quadtree_and_bounding_boxes(quadtree, bounding_boxes, x_min, x_max,
 * y_min, y_max, scale, max_depth) */
inline bbox_quad_pairs join_quadtree_and_bounding_boxes(point_quadtree const& quadtree,
                                                        bounding_boxes const& boxes, double x_min,
                                                        double x_max, double y_min, double y_max,
                                                        double scale, int8_t max_depth,
                                                        void* stream = nullptr,
                                                        qpip_resource const* mr = nullptr)
{
  uint32_t *bo = nullptr, *qo = nullptr;
  int64_t k = 0;
  detail::check(qpip_join_quadtree_and_bounding_boxes(
    quadtree.key.data(), quadtree.level.data(), quadtree.is_internal_node.data(),
    quadtree.length.data(), quadtree.offset.data(), quadtree.size(), boxes.minx.data(),
    boxes.miny.data(), boxes.maxx.data(), boxes.maxy.data(), boxes.size(), x_min, x_max, y_min,
    y_max, scale, max_depth, mr, stream, &bo, &qo, &k));
  return bbox_quad_pairs{column<uint32_t>(bo, k, mr, stream), column<uint32_t>(qo, k, mr, stream)};
}

/* quadtree_point_in_polygon(poly_quad_pairs, quadtree, point_indices, point_x,
 * point_y, poly_offsets, ring_offsets, poly_points_x, poly_points_y) */
inline polygon_point_pairs quadtree_point_in_polygon(
  bbox_quad_pairs const& pairs, point_quadtree const& quadtree,
  column<uint32_t> const& point_indices, double const* x, int64_t num_x, double const* y,
  int64_t num_y, int32_t const* poly_offsets, int64_t num_poly_offsets,
  int32_t const* ring_offsets, int64_t num_ring_offsets, double const* poly_x, int64_t num_poly_x,
  double const* poly_y, int64_t num_poly_y, void* stream = nullptr,
  qpip_resource const* mr = nullptr)
{
  uint32_t *po = nullptr, *pt = nullptr;
  int64_t h = 0;
  detail::check(qpip_quadtree_point_in_polygon(
    pairs.bbox_offset.data(), pairs.quad_offset.data(), pairs.size(), quadtree.key.data(),
    quadtree.level.data(), quadtree.is_internal_node.data(), quadtree.length.data(),
    quadtree.offset.data(), quadtree.size(), point_indices.data(), point_indices.size(), x, y,
    num_x, num_y, reinterpret_cast<uint32_t const*>(poly_offsets), num_poly_offsets,
    reinterpret_cast<uint32_t const*>(ring_offsets), num_ring_offsets, poly_x, poly_y,
    num_poly_x, num_poly_y, mr, stream, &po, &pt, &h));
  return polygon_point_pairs{column<uint32_t>(po, h, mr, stream),
                             column<uint32_t>(pt, h, mr, stream)};
}

/* quadtree_point_to_nearest_linestring(linestring_quad_pairs, quadtree,
 * point_indices, point_x, point_y, linestring_offsets, linestring_points_x,
 * linestring_points_y) */
inline point_linestring_distances quadtree_point_to_nearest_linestring(
  bbox_quad_pairs const& pairs, point_quadtree const& quadtree,
  column<uint32_t> const& point_indices, double const* x, int64_t num_x, double const* y,
  int64_t num_y, uint32_t const* linestring_offsets, int64_t num_linestring_offsets,
  double const* linestring_x, int64_t num_linestring_x, double const* linestring_y,
  int64_t num_linestring_y, void* stream = nullptr, qpip_resource const* mr = nullptr)
{
  uint32_t *pi = nullptr, *li = nullptr;
  double* d = nullptr;
  int64_t n = 0;
  detail::check(qpip_quadtree_point_to_nearest_linestring(
    pairs.bbox_offset.data(), pairs.quad_offset.data(), pairs.size(), quadtree.key.data(),
    quadtree.level.data(), quadtree.is_internal_node.data(), quadtree.length.data(),
    quadtree.offset.data(), quadtree.size(), point_indices.data(), point_indices.size(), x, y,
    num_x, num_y, linestring_offsets, num_linestring_offsets, linestring_x, linestring_y,
    num_linestring_x, num_linestring_y, mr, stream, &pi, &li, &d, &n));
  return point_linestring_distances{column<uint32_t>(pi, n, mr, stream),
                                    column<uint32_t>(li, n, mr, stream),
                                    column<double>(d, n, mr, stream)};
}

}  // namespace cuspatial_amd

#endif /* CUSPATIAL_AMD_HPP */
