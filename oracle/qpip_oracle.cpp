// Oracle: a host C++ restatement of the reference quadtree point-in-polygon path.
//
// TEST INFRASTRUCTURE ONLY. This file is the parity checker and the CPU baseline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// The product library (cuspatial_amd/) never links, loads or calls it.
//
// Each function follows the reference step by step, in the same order and with
// the same intermediate arrays, so that a reader can check it line by line
// against the cited reference files (paths relative to the reference root).
//
// Arithmetic policy (both here and in the HIP kernels):
//  * FP64 throughout, no contraction (compiled with -ffp-contract=off) except
//    where the reference's compiled code fuses: the node-extent expression in
//    detail/join/intersection.cuh:116-119 (`v_min + key * level_scale`) is built
//    by nvcc with its default --fmad=true (cpp/cmake/Modules/ConfigureCUDA.cmake
//    sets no --fmad=false), so it is computed here with an explicit std::fma.
//  * float -> uint16 conversion of (x - min) / scale truncates toward zero; NaN
//    maps to 0 and values >= 65535 saturate (PTX cvt.rzi semantics). Points
//    inside the bounding box never reach the saturating range because the scale
//    is clamped to extent / (2^d + 2).
//
// Reference quirks kept on purpose (all pinned by source text only):
//  * the level loop drops the last reduction (the "root") and computes
//    num_parent_nodes = end - Q - 1 (phase_1.cuh:175-180, 211);
//  * points outside the bbox get the bottom-level key (1 << 2d) - 1 (phase_1.cuh:79-82);
//  * zero-length ring edges `continue` without updating b / y0
//    (detail/algorithm/is_point_in_polygon.cuh:65);
//  * the pruning stencil of remove_unqualified_quads (phase_2.cuh:270-281) aliases
//    the range being compacted. This restatement evaluates the stencil on the
//    pre-compaction point counts (the values count_if at :256-261 reads), which
//    is what the algorithm (Zhang et al., ADMS'19, Fig. 5 line 5) specifies;
//  * compute_parent_positions scatters parents with zero children to
//    position_map[i] == num_child_nodes, one past the end of parent_pos, when the
//    root level has more than one node (phase_2.cuh:204-208). Such writes are
//    skipped here (they cannot be reproduced, the reference writes out of bounds).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Result object handed across the C boundary.
// ---------------------------------------------------------------------------
struct Result {
  int status = 0;
  std::string error;
  std::vector<std::vector<uint8_t>> arrays;
  std::vector<int64_t> counts;

  template <class T>
  void add(std::vector<T> const& v)
  {
    std::vector<uint8_t> bytes(v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(bytes.data(), v.data(), bytes.size());
    arrays.push_back(std::move(bytes));
    counts.push_back(static_cast<int64_t>(v.size()));
  }
};

Result* fail(std::string const& msg)
{
  auto* r   = new Result;
  r->status = 1;
  r->error  = "cuSpatial failure: " + msg;
  return r;
}

// ---------------------------------------------------------------------------
// z-order (detail/utility/z_order.cuh:25-94). The reference spreads the bits of
// each byte through a 256-entry table; the closed form below spreads bit i of a
// 16-bit value to bit 2i.
// ---------------------------------------------------------------------------
inline uint32_t dilate_2(uint16_t v)
{
  uint32_t x = v;
  x          = (x | (x << 8)) & 0x00FF00FFu;
  x          = (x | (x << 4)) & 0x0F0F0F0Fu;
  x          = (x | (x << 2)) & 0x33333333u;
  x          = (x | (x << 1)) & 0x55555555u;
  return x;
}

inline uint32_t z_order(uint16_t x, uint16_t y) { return (dilate_2(y) << 1) | dilate_2(x); }

// z_order.cuh:82-94: undilate of (index & 0x55555555) and ((index >> 1) & 0x55555555).
inline uint16_t undilate_2(uint32_t x)
{
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0F0F0F0Fu;
  x = (x | (x >> 4)) & 0x00FF00FFu;
  x = (x | (x >> 8)) & 0x0000FFFFu;
  return static_cast<uint16_t>(x);
}
inline uint16_t z_order_x(uint32_t key) { return undilate_2(key & 0x55555555u); }
inline uint16_t z_order_y(uint32_t key) { return undilate_2((key >> 1) & 0x55555555u); }

// Box reduction operators (geometry/vec_2d.hpp:185-205: device ::min / ::max,
// which ignore a NaN operand). -0.0 orders before +0.0 so that the result does
// not depend on the order in which a parallel reduction meets the operands.
inline double box_min(double a, double b)
{
  if (std::isnan(a)) return b;
  if (std::isnan(b)) return a;
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}
inline double box_max(double a, double b)
{
  if (std::isnan(a)) return b;
  if (std::isnan(b)) return a;
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}

inline uint16_t to_u16(double v)
{
  if (!(v >= 0.0)) return 0;  // NaN and negatives -> 0
  if (v >= 65535.0) return 65535;
  return static_cast<uint16_t>(v);
}

// ---------------------------------------------------------------------------
// float_equal (detail/utility/floating_point.cuh:97-129): biased-integer ULP
// distance <= 4, NaN never equal.
// ---------------------------------------------------------------------------
inline uint64_t signmagnitude_to_biased(uint64_t sam)
{
  uint64_t const sign = uint64_t{1} << 63;
  return (sam & sign) ? (~sam + 1) : (sam | sign);
}

inline bool float_equal(double a, double b)
{
  if (std::isnan(a) || std::isnan(b)) return false;
  uint64_t ab, bb;
  std::memcpy(&ab, &a, 8);
  std::memcpy(&bb, &b, 8);
  uint64_t const la = signmagnitude_to_biased(ab);
  uint64_t const lb = signmagnitude_to_biased(bb);
  return la >= lb ? (la - lb) <= 4 : (lb - la) <= 4;
}

// ---------------------------------------------------------------------------
// is_point_in_polygon (detail/algorithm/is_point_in_polygon.cuh:46-100), planar.
// The polygon is rings [ring_begin, ring_end) of ring_offsets; ring r spans
// vertices [ring_offsets[r], ring_offsets[r+1]).
// ---------------------------------------------------------------------------
template <class Off>
bool is_point_in_polygon(double px,
                         double py,
                         Off const* ring_offsets,
                         int64_t ring_begin,
                         int64_t ring_end,
                         double const* vx,
                         double const* vy)
{
  bool point_is_within = false;
  bool point_on_edge   = false;
  for (int64_t r = ring_begin; r < ring_end; ++r) {
    int64_t const vb = static_cast<int64_t>(ring_offsets[r]);
    int64_t const ve = static_cast<int64_t>(ring_offsets[r + 1]);
    if (ve <= vb) continue;  // an empty ring contributes no segment
    // last_segment.v2 is the last vertex of the ring
    double bx    = vx[ve - 1];
    double by    = vy[ve - 1];
    bool y0_flag = by > py;
    bool y1_flag;
    for (int64_t v = vb; v < ve; ++v) {
      double const ax   = vx[v];
      double const ay   = vy[v];
      double const run  = bx - ax;
      double const rise = by - ay;
      if (float_equal(run, 0.0) && float_equal(rise, 0.0)) continue;
      double const rise_to_point = py - ay;
      double const run_to_point  = px - ax;
      bool const is_collinear    = float_equal(run * rise_to_point, run_to_point * rise);
      if (is_collinear) {
        double minx = ax;
        double maxx = bx;
        if (minx > maxx) std::swap(minx, maxx);
        if (minx <= px && px <= maxx) {
          point_on_edge = true;
          break;
        }
      }
      y1_flag = ay > py;
      if (y1_flag != y0_flag) {
        double const lhs = (px - ax) * rise;
        double const rhs = run * rise_to_point;
        if ((lhs < rhs) != y1_flag) { point_is_within = !point_is_within; }
      }
      bx      = ax;
      by      = ay;
      y0_flag = y1_flag;
    }
    if (point_on_edge) {
      point_is_within = false;
      break;
    }
  }
  return point_is_within;
}

// ---------------------------------------------------------------------------
// Quadtree construction.
// ---------------------------------------------------------------------------
struct Tree {
  std::vector<uint32_t> key;
  std::vector<uint8_t> level;
  std::vector<uint8_t> is_internal;
  std::vector<uint32_t> length;
  std::vector<uint32_t> offset;
};

// thrust::reduce_by_key with a unary count value (phase_1.cuh:108-125 at :279-285)
int64_t reduce_keys_count(std::vector<uint32_t> const& keys,
                          std::vector<uint32_t>& out_keys,
                          std::vector<uint32_t>& out_count)
{
  int64_t q = 0;
  for (size_t i = 0; i < keys.size(); ++i) {
    if (i == 0 || keys[i] != keys[i - 1]) {
      out_keys[q]  = keys[i];
      out_count[q] = 0;
      ++q;
    }
    out_count[q - 1] += 1;
  }
  return q;
}

// phase_1.cuh:134-181 build_tree_levels. Arrays are the Q*(d+1)-sized buffers.
int64_t build_tree_levels(std::vector<uint32_t>& keys,
                          std::vector<uint32_t>& pcount,
                          std::vector<uint32_t>& ccount,
                          int max_depth,
                          int64_t num_top_quads,
                          int64_t& end_out,
                          std::vector<int64_t>& begin_pos,
                          std::vector<int64_t>& end_pos)
{
  int64_t begin = 0;
  int64_t end   = num_top_quads;
  begin_pos.assign(max_depth, 0);
  end_pos.assign(max_depth, 0);
  for (int level = max_depth - 1; level >= 0; --level) {
    // reduce_by_key(parent_keys = key >> 2, values = (pcount, 1), tuple_sum)
    int64_t n_out = 0;
    for (int64_t i = begin; i < end; ++i) {
      uint32_t const pk = keys[i] >> 2;
      if (i == begin || pk != (keys[i - 1] >> 2)) {
        keys[end + n_out]   = pk;
        pcount[end + n_out] = 0;
        ccount[end + n_out] = 0;
        ++n_out;
      }
      pcount[end + n_out - 1] += pcount[i];
      ccount[end + n_out - 1] += 1;
    }
    end_pos[level]   = end;
    begin_pos[level] = begin;
    begin            = end;
    end              = end + n_out;
  }
  end_out = end;
  return end - num_top_quads - 1;
}

Tree construct_quadtree(double const* xs,
                        double const* ys,
                        int64_t n,
                        double min_x,
                        double min_y,
                        double max_x,
                        double max_y,
                        double scale,
                        int max_depth,
                        int32_t max_size,
                        std::vector<uint32_t>& point_indices)
{
  // --- compute_point_keys_and_sorted_indices (phase_1.cuh:62-95)
  std::vector<uint32_t> keys(n);
  uint32_t const outside_key = static_cast<uint32_t>((1u << (2 * max_depth)) - 1u);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double const x = xs[i], y = ys[i];
    if (x < min_x || x > max_x || y < min_y || y > max_y) {
      keys[i] = outside_key;
    } else {
      keys[i] = z_order(to_u16((x - min_x) / scale), to_u16((y - min_y) / scale));
    }
  }
  point_indices.resize(n);
  std::iota(point_indices.begin(), point_indices.end(), 0u);
  // thrust::stable_sort_by_key(keys, indices)
  std::stable_sort(point_indices.begin(), point_indices.end(), [&](uint32_t a, uint32_t b) {
    return keys[a] < keys[b];
  });
  {
    std::vector<uint32_t> sorted(n);
    for (int64_t i = 0; i < n; ++i) sorted[i] = keys[point_indices[i]];
    keys.swap(sorted);
  }

  // --- make_full_levels (phase_1.cuh:256-381)
  std::vector<uint32_t> quad_keys(n), quad_point_count(n);
  int64_t const num_bottom_quads = reduce_keys_count(keys, quad_keys, quad_point_count);
  size_t const full_size         = static_cast<size_t>(num_bottom_quads) * (max_depth + 1);
  quad_keys.resize(full_size);
  quad_point_count.resize(full_size);
  std::vector<uint32_t> quad_child_count(full_size, 0);

  int64_t quad_tree_size = 0;
  std::vector<int64_t> begin_pos, end_pos;
  int64_t num_parent_nodes = build_tree_levels(quad_keys,
                                               quad_point_count,
                                               quad_child_count,
                                               max_depth,
                                               num_bottom_quads,
                                               quad_tree_size,
                                               begin_pos,
                                               end_pos);
  quad_keys.resize(quad_tree_size);
  quad_point_count.resize(quad_tree_size);
  quad_child_count.resize(quad_tree_size);

  Tree t;
  if (num_parent_nodes <= 0) {
    // make_leaf_tree (detail/point_quadtree.cuh:155-188)
    int64_t const q = num_bottom_quads;
    t.key.assign(quad_keys.begin(), quad_keys.begin() + q);
    t.length.assign(quad_point_count.begin(), quad_point_count.begin() + q);
    t.level.assign(q, 0);
    t.is_internal.assign(q, 0);
    t.offset.resize(q);
    uint32_t acc = 0;
    for (int64_t i = 0; i < q; ++i) {
      t.offset[i] = acc;
      acc += t.length[i];
    }
    return t;
  }

  // reverse_tree_levels (phase_1.cuh:197-238); the tail past the copied levels
  // stays zero here (uninitialised in the reference, never read).
  std::vector<uint32_t> keys_r(quad_tree_size, 0), pc_r(quad_tree_size, 0), cc_r(quad_tree_size, 0);
  std::vector<uint8_t> lev_r(quad_tree_size, 0);
  int64_t off = 0;
  for (int level = 0; level < max_depth; ++level) {
    for (int64_t i = begin_pos[level]; i < end_pos[level]; ++i, ++off) {
      keys_r[off] = quad_keys[i];
      pc_r[off]   = quad_point_count[i];
      cc_r[off]   = quad_child_count[i];
      lev_r[off]  = static_cast<uint8_t>(level);
    }
  }
  int64_t const level_1_size = end_pos[0] - begin_pos[0];

  // --- make_quad_tree (detail/point_quadtree.cuh:43-150)
  int64_t num_child_nodes = 0;
  for (int64_t i = 0; i < num_parent_nodes; ++i) num_child_nodes += cc_r[i];

  // remove_unqualified_quads (phase_2.cuh:236-296)
  //   compute_parent_positions (phase_2.cuh:186-220)
  std::vector<uint32_t> parent_pos(num_child_nodes, 0);
  {
    uint32_t acc = 0;
    for (int64_t i = 0; i < num_parent_nodes; ++i) {
      uint32_t const pos = acc;  // exclusive scan of child counts
      acc += cc_r[i];
      if (static_cast<int64_t>(pos) < num_child_nodes) parent_pos[pos] = static_cast<uint32_t>(i);
    }
    uint32_t run = 0;
    for (int64_t j = 0; j < num_child_nodes; ++j) {
      run           = std::max(run, parent_pos[j]);
      parent_pos[j] = run;
    }
  }
  // stencil snapshot: parent point counts before compaction
  std::vector<uint32_t> parent_pc(num_child_nodes);
  for (int64_t j = 0; j < num_child_nodes; ++j) parent_pc[j] = pc_r[parent_pos[j]];

  int64_t num_invalid_parent_nodes = 0;
  for (int64_t j = 0; j < num_parent_nodes - level_1_size; ++j) {
    if (parent_pc[j] <= static_cast<uint32_t>(max_size)) ++num_invalid_parent_nodes;
  }
  int64_t w = level_1_size;
  for (int64_t j = 0; j < num_child_nodes; ++j) {
    if (parent_pc[j] <= static_cast<uint32_t>(max_size)) continue;
    int64_t const src = level_1_size + j;
    keys_r[w]         = keys_r[src];
    pc_r[w]           = pc_r[src];
    cc_r[w]           = cc_r[src];
    lev_r[w]          = lev_r[src];
    ++w;
  }
  int64_t const num_valid_nodes = w;
  keys_r.resize(num_valid_nodes);
  pc_r.resize(num_valid_nodes);
  cc_r.resize(num_valid_nodes);
  lev_r.resize(num_valid_nodes);

  num_parent_nodes -= num_invalid_parent_nodes;

  // construct_non_leaf_indicator (phase_2.cuh:308-345)
  std::vector<uint8_t> is_internal(num_valid_nodes, 0);
  for (int64_t i = 0; i < num_parent_nodes && i < num_valid_nodes; ++i) {
    is_internal[i] = pc_r[i] > static_cast<uint32_t>(max_size);
    if (is_internal[i]) pc_r[i] = 0;
  }

  // compute_flattened_first_point_positions (phase_2.cuh:105-184)
  std::vector<uint32_t> quad_point_pos(num_valid_nodes, 0);
  {
    int const last_level = max_depth - 1;
    std::vector<uint32_t> flattened(num_valid_nodes);
    for (int64_t i = 0; i < num_valid_nodes; ++i) {
      flattened[i] = is_internal[i] ? 0xFFFFFFFFu : (keys_r[i] << (2 * (last_level - lev_r[i])));
    }
    std::vector<uint32_t> order(num_valid_nodes);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return flattened[a] < flattened[b];
    });
    std::vector<uint32_t> pc_tmp;  // remove_if(pc == 0) over the sorted counts
    for (int64_t i = 0; i < num_valid_nodes; ++i) {
      if (pc_r[order[i]] != 0) pc_tmp.push_back(pc_r[order[i]]);
    }
    std::vector<uint32_t> leaf_pos;  // copy_if(!is_internal) of node indices
    for (int64_t i = 0; i < num_valid_nodes; ++i) {
      if (!is_internal[i]) leaf_pos.push_back(static_cast<uint32_t>(i));
    }
    size_t const nleaf = leaf_pos.size();
    pc_tmp.resize(nleaf, 0);
    order.resize(nleaf);
    std::vector<uint32_t> offs_tmp(nleaf);
    uint32_t acc = 0;
    for (size_t i = 0; i < nleaf; ++i) {
      offs_tmp[i] = acc;
      acc += pc_tmp[i];
    }
    // stable_sort_by_key(order, (pc_tmp, offs_tmp))
    std::vector<size_t> perm(nleaf);
    std::iota(perm.begin(), perm.end(), size_t{0});
    std::stable_sort(
      perm.begin(), perm.end(), [&](size_t a, size_t b) { return order[a] < order[b]; });
    for (size_t i = 0; i < nleaf; ++i) {
      pc_r[leaf_pos[i]]           = pc_tmp[perm[i]];
      quad_point_pos[leaf_pos[i]] = offs_tmp[perm[i]];
    }
  }

  // make_quad_tree lines 9-10 + selects (detail/point_quadtree.cuh:88-133)
  for (int64_t i = 0; i < num_valid_nodes; ++i) {
    if (!is_internal[i]) cc_r[i] = 0;
  }
  t.key    = keys_r;
  t.level  = lev_r;
  t.is_internal = is_internal;
  t.length.resize(num_valid_nodes);
  t.offset.resize(num_valid_nodes);
  uint32_t child_pos = static_cast<uint32_t>(level_1_size);
  for (int64_t i = 0; i < num_valid_nodes; ++i) {
    uint32_t const cp = child_pos;
    child_pos += cc_r[i];
    t.offset[i] = is_internal[i] ? cp : quad_point_pos[i];
    t.length[i] = is_internal[i] ? cc_r[i] : pc_r[i];
  }
  return t;
}

// ---------------------------------------------------------------------------
// Join (detail/join/intersection.cuh, traversal.cuh, quadtree_bbox_filtering.cuh)
// ---------------------------------------------------------------------------
struct PairList {
  std::vector<uint32_t> node;
  std::vector<uint32_t> bbox;
};

// find_intersections' per-pair functor: returns 0 leaf, 1 quad, 2 none.
inline uint8_t classify(uint32_t key,
                        uint8_t level,
                        uint8_t is_internal,
                        double bminx,
                        double bminy,
                        double bmaxx,
                        double bmaxy,
                        double vminx,
                        double vminy,
                        double scale,
                        int max_depth)
{
  double const key_x = static_cast<double>(z_order_x(key));
  double const key_y = static_cast<double>(z_order_y(key));
  int const shift    = max_depth - 1 - static_cast<int>(level);
  double const level_scale = scale * static_cast<double>(1 << (shift < 0 ? 0 : shift));
  double const node_x_min  = std::fma(key_x + 0, level_scale, vminx);
  double const node_y_min  = std::fma(key_y + 0, level_scale, vminy);
  double const node_x_max  = std::fma(key_x + 1, level_scale, vminx);
  double const node_y_max  = std::fma(key_y + 1, level_scale, vminy);
  if ((node_x_min > bmaxx) || (node_x_max < bminx) || (node_y_min > bmaxy) ||
      (node_y_max < bminy)) {
    return 2;
  }
  return is_internal ? 1 : 0;
}

}  // namespace

extern "C" {

// ---- result accessors ------------------------------------------------------
int oq_result_status(void* h) { return static_cast<Result*>(h)->status; }
const char* oq_result_error(void* h) { return static_cast<Result*>(h)->error.c_str(); }
int oq_result_num_arrays(void* h) { return static_cast<int>(static_cast<Result*>(h)->arrays.size()); }
int64_t oq_result_count(void* h, int i) { return static_cast<Result*>(h)->counts.at(i); }
const void* oq_result_data(void* h, int i) { return static_cast<Result*>(h)->arrays.at(i).data(); }
void oq_result_free(void* h) { delete static_cast<Result*>(h); }

void oq_set_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n < 1 ? 1 : n);
#else
  (void)n;
#endif
}

int oq_max_threads()
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---- primitives for known-answer tests ------------------------------------
uint32_t oq_z_order(uint16_t x, uint16_t y) { return z_order(x, y); }
uint16_t oq_z_order_x(uint32_t k) { return z_order_x(k); }
uint16_t oq_z_order_y(uint32_t k) { return z_order_y(k); }
int oq_float_equal(double a, double b) { return float_equal(a, b) ? 1 : 0; }

int oq_is_point_in_polygon(double px,
                           double py,
                           const uint32_t* ring_offsets,
                           int64_t ring_begin,
                           int64_t ring_end,
                           const double* vx,
                           const double* vy)
{
  return is_point_in_polygon(px, py, ring_offsets, ring_begin, ring_end, vx, vy) ? 1 : 0;
}

// Brute-force bitmask PIP of points against <= 31 polygons, the reference's
// large-test oracle (detail/point_in_polygon.cuh:42-66): bit j of out[i] is set
// when point i is inside polygon j (polygon j = rings [poly_offsets[j], poly_offsets[j+1])).
void oq_point_in_polygon_bitmask(const double* x,
                                 const double* y,
                                 int64_t n,
                                 const uint32_t* poly_offsets,
                                 int64_t n_poly_offsets,
                                 const uint32_t* ring_offsets,
                                 const double* vx,
                                 const double* vy,
                                 int32_t* out)
{
  int64_t const npoly = n_poly_offsets - 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int32_t bits = 0;
    for (int64_t j = 0; j < npoly; ++j) {
      if (is_point_in_polygon(x[i], y[i], ring_offsets, poly_offsets[j], poly_offsets[j + 1], vx, vy))
        bits |= (1 << j);
    }
    out[i] = bits;
  }
}

// pairwise_point_in_polygon (detail/point_in_polygon.cuh:104-160 with
// pairwise_point_in_polygon_functor): point i against polygon i, 0 or 1.
void oq_pairwise_point_in_polygon(const double* x,
                                  const double* y,
                                  int64_t n,
                                  const uint32_t* poly_offsets,
                                  const uint32_t* ring_offsets,
                                  const double* vx,
                                  const double* vy,
                                  uint8_t* out)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    out[i] = is_point_in_polygon(x[i], y[i], ring_offsets, poly_offsets[i], poly_offsets[i + 1],
                                 vx, vy)
               ? 1
               : 0;
}

// All (polygon, point) pairs with the point strictly inside, by brute force over
// every polygon and every point, ordered by point then polygon. Used as the
// set-equality check of the reference's large test (quadtree_point_in_polygon_test_large.cu:170-224).
void* oq_all_pairs_bruteforce(const double* x,
                              const double* y,
                              int64_t n,
                              const uint32_t* poly_offsets,
                              int64_t n_poly_offsets,
                              const uint32_t* ring_offsets,
                              const double* vx,
                              const double* vy)
{
  int64_t const npoly = n_poly_offsets - 1;
  std::vector<std::vector<uint32_t>> per_point(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t j = 0; j < npoly; ++j) {
      if (is_point_in_polygon(x[i], y[i], ring_offsets, poly_offsets[j], poly_offsets[j + 1], vx, vy))
        per_point[i].push_back(static_cast<uint32_t>(j));
    }
  }
  std::vector<uint32_t> polys, points;
  for (int64_t i = 0; i < n; ++i) {
    for (auto j : per_point[i]) {
      polys.push_back(j);
      points.push_back(static_cast<uint32_t>(i));
    }
  }
  auto* r = new Result;
  r->add(polys);
  r->add(points);
  return r;
}

// ---- quadtree_on_points ----------------------------------------------------
// cpp/src/indexing/point_quadtree.cu:154-180 (checks, empty) and
// detail/point_quadtree.cuh:238-272 (clamps). Arrays: point_indices, key, level,
// is_internal_node, length, offset.
void* oq_quadtree_on_points(const double* x,
                            const double* y,
                            int64_t nx,
                            int64_t ny,
                            double x_min,
                            double x_max,
                            double y_min,
                            double y_max,
                            double scale,
                            int32_t max_depth_in,
                            int32_t max_size)
{
  if (nx != ny) return fail("x and y columns must have the same length");
  auto* r = new Result;
  std::vector<uint32_t> point_indices;
  Tree t;
  if (nx > 0) {
    double const min_x = std::min(x_min, x_max), max_x = std::max(x_min, x_max);
    double const min_y = std::min(y_min, y_max), max_y = std::max(y_min, y_max);
    max_size           = std::max(1, max_size);
    int8_t md          = static_cast<int8_t>(max_depth_in);
    md                 = std::max(int8_t{0}, std::min(int8_t{15}, md));
    int const d        = md;
    scale = std::max(scale, std::max(max_x - min_x, max_y - min_y) / ((1 << d) + 2));
    t = construct_quadtree(x, y, nx, min_x, min_y, max_x, max_y, scale, d, max_size, point_indices);
  }
  r->add(point_indices);
  r->add(t.key);
  r->add(t.level);
  r->add(t.is_internal);
  r->add(t.length);
  r->add(t.offset);
  return r;
}

// ---- polygon_bounding_boxes ------------------------------------------------
// cpp/src/bounding_boxes/polygon_bounding_boxes.cu:132-159 and
// detail/bounding_boxes.cuh:62-90,139-184: reduce_by_key of per-vertex boxes keyed
// by the polygon id of each vertex (upper_bound over the first ring offset of
// each polygon, iterator_factory.cuh:168-182,415-426). Runs of equal ids are
// written in order, so an empty polygon produces no box and later boxes move up;
// unwritten trailing entries are NaN here (uninitialised in the reference).
void* oq_polygon_bounding_boxes(const int32_t* poly_offsets,
                                int64_t n_poly_offsets,
                                const int32_t* ring_offsets,
                                int64_t n_ring_offsets,
                                const double* x,
                                const double* y,
                                int64_t nx,
                                int64_t ny,
                                double expansion_radius)
{
  if (nx != ny) return fail("x and y must be the same size");
  int64_t const num_polys    = n_poly_offsets > 0 ? n_poly_offsets - 1 : 0;
  int64_t const num_rings    = n_ring_offsets > 0 ? n_ring_offsets - 1 : 0;
  double const nan           = std::nan("");
  std::vector<double> minx(num_polys, nan), miny(num_polys, nan), maxx(num_polys, nan),
    maxy(num_polys, nan);
  if (num_polys > 0 && num_rings > 0 && nx > 0) {
    // geometry ids: first ring offset of each polygon
    std::vector<int64_t> first(n_poly_offsets);
    for (int64_t i = 0; i < n_poly_offsets; ++i) first[i] = ring_offsets[poly_offsets[i]];
    int64_t out = -1;
    int64_t prev_id = -1;
    for (int64_t v = 0; v < nx; ++v) {
      int64_t const id = std::upper_bound(first.begin(), first.end(), v) - first.begin();
      double const lx = x[v] - expansion_radius, hx = x[v] + expansion_radius;
      double const ly = y[v] - expansion_radius, hy = y[v] + expansion_radius;
      if (id != prev_id) {
        ++out;
        prev_id = id;
        if (out >= num_polys) break;  // the reference would write past the output
        minx[out] = lx;
        miny[out] = ly;
        maxx[out] = hx;
        maxy[out] = hy;
      } else {
        minx[out] = box_min(minx[out], lx);
        miny[out] = box_min(miny[out], ly);
        maxx[out] = box_max(maxx[out], hx);
        maxy[out] = box_max(maxy[out], hy);
      }
    }
  }
  auto* r = new Result;
  r->add(minx);
  r->add(miny);
  r->add(maxx);
  r->add(maxy);
  return r;
}

// ---- join_quadtree_and_bounding_boxes ---------------------------------------
// cpp/src/join/quadtree_bbox_filtering.cu:90-126 (checks, empty) and
// detail/join/quadtree_bbox_filtering.cuh:37-188 (level-synchronous BFS, final
// stable sort by the leaf's offset). Arrays: bbox_offset, quad_offset.
void* oq_join_quadtree_and_bounding_boxes(const uint32_t* key,
                                          const uint8_t* level,
                                          const uint8_t* is_internal,
                                          const uint32_t* length,
                                          const uint32_t* offset,
                                          int64_t m,
                                          const double* bminx,
                                          const double* bminy,
                                          const double* bmaxx,
                                          const double* bmaxy,
                                          int64_t nb,
                                          double x_min,
                                          double x_max,
                                          double y_min,
                                          double y_max,
                                          double scale,
                                          int32_t max_depth)
{
  if (!(scale > 0)) return fail("scale must be positive");
  if (!(x_min < x_max && y_min < y_max))
    return fail("invalid bounding box (x_min, x_max, y_min, y_max)");
  if (!(max_depth > 0 && max_depth < 16))
    return fail("maximum depth must be positive and less than 16");
  auto* r = new Result;
  std::vector<uint32_t> out_bbox, out_node;
  if (m > 0 && nb > 0) {
    int64_t num_top = 0;
    for (int64_t i = 0; i < m; ++i) num_top += (level[i] == 0);
    auto test = [&](uint32_t node, uint32_t bbox) {
      return classify(key[node], level[node], is_internal[node], bminx[bbox], bminy[bbox],
                      bmaxx[bbox], bmaxy[bbox], x_min, y_min, scale, max_depth);
    };
    PairList cur;
    if (num_top > 0) {
      for (int64_t i = 0; i < num_top * nb; ++i) {
        uint32_t const node = static_cast<uint32_t>(i % num_top);
        uint32_t const bbox = static_cast<uint32_t>(i / num_top);
        uint8_t const t     = test(node, bbox);
        if (t == 0) {
          out_node.push_back(node);
          out_bbox.push_back(bbox);
        } else if (t == 1) {
          cur.node.push_back(node);
          cur.bbox.push_back(bbox);
        }
      }
    }
    for (int lvl = 1; lvl < max_depth && !cur.node.empty(); ++lvl) {
      PairList next;
      for (size_t p = 0; p < cur.node.size(); ++p) {
        uint32_t const parent = cur.node[p];
        for (uint32_t j = 0; j < length[parent]; ++j) {
          uint32_t const child = offset[parent] + j;
          uint8_t const t      = test(child, cur.bbox[p]);
          if (t == 0) {
            out_node.push_back(child);
            out_bbox.push_back(cur.bbox[p]);
          } else if (t == 1) {
            next.node.push_back(child);
            next.bbox.push_back(cur.bbox[p]);
          }
        }
      }
      cur.node.swap(next.node);
      cur.bbox.swap(next.bbox);
    }
    // stable_sort_by_key(offset[node], (bbox, node))
    std::vector<size_t> perm(out_node.size());
    std::iota(perm.begin(), perm.end(), size_t{0});
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) {
      return offset[out_node[a]] < offset[out_node[b]];
    });
    std::vector<uint32_t> sb(perm.size()), sn(perm.size());
    for (size_t i = 0; i < perm.size(); ++i) {
      sb[i] = out_bbox[perm[i]];
      sn[i] = out_node[perm[i]];
    }
    out_bbox.swap(sb);
    out_node.swap(sn);
  }
  r->add(out_bbox);
  r->add(out_node);
  return r;
}

// ---- quadtree_point_in_polygon ----------------------------------------------
// cpp/src/join/quadtree_point_in_polygon.cu:143-191 (checks, empty) and
// detail/join/quadtree_point_in_polygon.cuh:104-235: every point of every
// (polygon, leaf) pair is tested; output in pair order then point order.
// Arrays: polygon_index, point_index (point_index is a position in point_indices).
void* oq_quadtree_point_in_polygon(const uint32_t* pair_poly,
                                   const uint32_t* pair_quad,
                                   int64_t k,
                                   const uint32_t* length,
                                   const uint32_t* offset,
                                   int64_t m,
                                   const uint32_t* point_indices,
                                   int64_t n_point_indices,
                                   const double* x,
                                   const double* y,
                                   int64_t nx,
                                   int64_t ny,
                                   const uint32_t* poly_offsets,
                                   int64_t n_poly_offsets,
                                   const uint32_t* ring_offsets,
                                   int64_t n_ring_offsets,
                                   const double* vx,
                                   const double* vy,
                                   int64_t nvx,
                                   int64_t nvy)
{
  (void)n_ring_offsets;
  if (!(n_point_indices == nx && nx == ny))
    return fail("number of points must be the same for both x and y columns");
  if (nvx != nvy) return fail("numbers of vertices must be the same for both x and y columns");
  auto* r = new Result;
  std::vector<uint32_t> out_poly, out_point;
  if (!(k == 0 || m == 0 || n_point_indices == 0 || n_poly_offsets == 0)) {
    std::vector<std::vector<uint32_t>> hits(k);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < k; ++p) {
      uint32_t const poly = pair_poly[p];
      uint32_t const quad = pair_quad[p];
      for (uint32_t local = 0; local < length[quad]; ++local) {
        uint32_t const pos = offset[quad] + local;
        uint32_t const idx = point_indices[pos];
        if (is_point_in_polygon(
              x[idx], y[idx], ring_offsets, poly_offsets[poly], poly_offsets[poly + 1], vx, vy))
          hits[p].push_back(pos);
      }
    }
    for (int64_t p = 0; p < k; ++p) {
      for (auto pos : hits[p]) {
        out_poly.push_back(pair_poly[p]);
        out_point.push_back(pos);
      }
    }
  }
  r->add(out_poly);
  r->add(out_point);
  return r;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// quadtree_point_to_nearest_linestring
// (detail/join/quadtree_point_to_nearest_linestring.cuh:143-314, step by step)
// ---------------------------------------------------------------------------
namespace {

// dot(a, b) = a.x * b.x + a.y * b.y contracted to fma(a.x, b.x, a.y * b.y): the
// reference's expected distances (quadtree_point_to_nearest_linestring_test_small.cu:
// 173-192) are reproduced bit for bit by this form and missed by up to 45 ULP by
// the unfused one.
double dot2(double ax, double ay, double bx, double by) { return std::fma(ax, bx, ay * by); }

// point_linestring_distance (detail/algorithm/point_linestring_distance.cuh:33-51).
double point_linestring_distance(double px, double py, const double* lx, const double* ly,
                                 uint32_t v0, uint32_t v1)
{
  double dsq = std::numeric_limits<double>::max();
  for (uint32_t i = v0; i + 1 < v1; ++i) {
    double const v1x = px - lx[i], v1y = py - ly[i];
    double const v2x = px - lx[i + 1], v2y = py - ly[i + 1];
    double const ex = lx[i + 1] - lx[i], ey = ly[i + 1] - ly[i];
    double const d0 = dot2(v1x, v1y, v1x, v1y);
    double const d1 = dot2(v2x, v2y, v2x, v2y);
    double const d2 = dot2(ex, ey, ex, ey);
    double const d3 = dot2(v1x, v1y, ex, ey);  // proj2: dot(point - v1, v2 - v1)
    double const r  = d3 * d3 / d2;
    double const d  = (d3 <= 0 || r >= d2) ? std::fmin(d0, d1) : d0 - r;
    dsq             = std::fmin(dsq, d);
  }
  return std::sqrt(dsq);
}

struct PLD {
  uint32_t point, linestring;
  double distance;
};

// The reduce_by_key operator (:279-294).
PLD nearest_op(PLD const& lhs, PLD const& rhs)
{
  if (lhs.distance == 0) return rhs;
  if (rhs.distance == 0) return lhs;
  if (lhs.distance == rhs.distance) return lhs.linestring < rhs.linestring ? lhs : rhs;
  return lhs.distance < rhs.distance ? lhs : rhs;
}

}  // namespace

extern "C" {

void* oq_quadtree_point_to_nearest_linestring(const uint32_t* pair_ls,
                                              const uint32_t* pair_quad,
                                              int64_t k,
                                              const uint32_t* length,
                                              const uint32_t* offset,
                                              int64_t m,
                                              const uint32_t* point_indices,
                                              int64_t n_point_indices,
                                              const double* x,
                                              const double* y,
                                              int64_t nx,
                                              int64_t ny,
                                              const uint32_t* ls_offsets,
                                              int64_t n_ls_offsets,
                                              const double* lx,
                                              const double* ly,
                                              int64_t nlx,
                                              int64_t nly)
{
  // cpp/src/join/quadtree_point_to_nearest_linestring.cu:153-176
  if (!(n_point_indices == nx && nx == ny))
    return fail("number of points must be the same for both x and y columns");
  if (nlx != nly) return fail("numbers of vertices must be the same for both x and y columns");
  auto* r = new Result;
  std::vector<uint32_t> out_point, out_ls;
  std::vector<double> out_dist;
  if (!(k == 0 || m == 0 || n_point_indices == 0 || n_ls_offsets == 0)) {
    // local point offsets: inclusive scan of the paired quadrants' lengths (:204-217)
    std::vector<uint64_t> offs(k + 1, 0);
    for (int64_t i = 0; i < k; ++i) offs[i + 1] = offs[i] + length[pair_quad[i]];
    uint64_t const total = offs[k];
    std::vector<uint32_t> quad_off(k);
    for (int64_t i = 0; i < k; ++i) quad_off[i] = offset[pair_quad[i]];
    // the transposed enumeration (:57-91, 106-118): within a run of pairs on one
    // quadrant, entries go point by point, each point against the run's linestrings
    std::vector<PLD> seq(total);
    for (uint64_t g = 0; g < total; ++g) {
      int64_t const pair =
        std::upper_bound(offs.begin(), offs.end(), g) - offs.begin() - 1;  // quad-linestring index
      uint64_t const local_point = g - offs[pair];
      uint32_t const qo          = quad_off[pair];
      int64_t const lhs = std::lower_bound(quad_off.begin(), quad_off.begin() + pair, qo) -
                          quad_off.begin();
      int64_t const rhs =
        std::upper_bound(quad_off.begin() + pair, quad_off.end(), qo) - quad_off.begin();
      uint64_t const local_ls   = static_cast<uint64_t>(pair - lhs);
      uint64_t const ls_in_quad = static_cast<uint64_t>(rhs - lhs);
      uint64_t const npts       = length[pair_quad[pair]];
      uint64_t const start      = local_ls * npts + local_point;
      uint32_t const point_id   = static_cast<uint32_t>(start / ls_in_quad + qo);
      int64_t const ls_pair     = static_cast<int64_t>(start % ls_in_quad) + lhs;
      uint32_t const ls         = pair_ls[ls_pair];
      uint32_t const idx        = point_indices[point_id];
      seq[g] = PLD{point_id, ls,
                   point_linestring_distance(x[idx], y[idx], lx, ly, ls_offsets[ls],
                                             ls_offsets[ls + 1])};
    }
    // reduce_by_key over equal consecutive point ids, then scatter to row point_id
    // (:266-312); rows never written: index columns left unset by the reference,
    // (p, 0xffffffff, 0) here
    out_point.resize(n_point_indices);
    out_ls.assign(n_point_indices, 0xffffffffu);
    out_dist.assign(n_point_indices, 0.0);
    for (int64_t p = 0; p < n_point_indices; ++p) out_point[p] = static_cast<uint32_t>(p);
    uint64_t g = 0;
    while (g < total) {
      PLD acc     = seq[g];
      uint64_t e  = g + 1;
      while (e < total && seq[e].point == acc.point) acc = nearest_op(acc, seq[e++]);
      out_point[acc.point] = acc.point;
      out_ls[acc.point]    = acc.linestring;
      out_dist[acc.point]  = acc.distance;
      g                    = e;
    }
  }
  r->add(out_point);
  r->add(out_ls);
  r->add(out_dist);
  return r;
}

}  // extern "C"
