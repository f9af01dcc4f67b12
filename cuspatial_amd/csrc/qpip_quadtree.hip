// quadtree_on_points on gfx950.
//
// Reference: cpp/src/indexing/point_quadtree.cu:154-180 (column API),
// detail/point_quadtree.cuh:238-272 (clamps), detail/index/construction/
// phase_1.cuh and phase_2.cuh (construction).
//
// The reference builds the tree bottom-up: it sorts the Morton keys, reduces the
// bottom cells, reduces every level above them into Q*(d+1)-sized arrays, then
// prunes the children of small nodes and recovers leaf offsets with two more
// sorts. This implementation computes the same five columns top-down from the
// sorted keys and never materialises pruned nodes:
//
//   1. keygen (fused with the radix sort's first digit count);
//   2. stable LSD radix sort of (key, point index) over the 2d+2 (at most 32)
//      key bits that can be non-zero;
//   3. level 0: the distinct values of key >> 2(d-1), found by binary search;
//   4. level l -> l+1: every node with more than max_size points and l <= d-2
//      (exactly the reference's internal nodes) is split into its non-empty
//      child quadrants by three binary searches inside its sorted point range;
//      children are appended in (parent, quadrant) order, i.e. level-major and
//      key order within a level - the order of reverse_tree_levels;
//   5. finalize: internal nodes get (length, offset) = (child count, first child
//      position) and leaves (point count, first point position).
//
// Equivalences with the reference (for the 5 output columns):
//   * a level-l node is kept by remove_unqualified_quads iff its parent has more
//     than max_size points (phase_2.cuh:256-281); since point counts shrink
//     downwards, that is iff all its ancestors are internal - the nodes generated
//     in step 4;
//   * a leaf's offset is the exclusive scan of leaf point counts in Morton order
//     (phase_2.cuh:105-184); leaves tile the sorted points in Morton order, so it
//     is the leaf's first sorted point position;
//   * when the root level holds R > 1 nodes (possible only if a point's cell
//     coordinate reaches 2^d, i.e. with a clamped scale), the reference's
//     num_parent_nodes = end - Q - 1 (phase_1.cuh:177) counts R - 1 extra
//     "parents", which are the first R - 1 bottom-level nodes in key order. Those
//     that survive pruning are flagged internal when they hold more than
//     max_size points, with length 0 and offset = number of nodes, and their
//     points drop out of the leaf-offset scan. Step 5 reproduces exactly that.
#include <algorithm>
#include <vector>

#include "qpip_sort.hpp"
#include "qpip_tree.hpp"

namespace qpip {

namespace {

constexpr int kMaxLevels = 16;
constexpr int kMaxMarked = 16;

struct TreeMeta {
  uint64_t sizes[kMaxLevels + 1];  // nodes per level (sizes[l] for l < d)
  uint32_t overflow;
  uint32_t num_marked;             // E = R - 1 (d >= 1)
  uint32_t marked_key[kMaxMarked];
  uint32_t marked_internal[kMaxMarked];
  uint32_t marked_start[kMaxMarked];
  uint32_t marked_count[kMaxMarked];
  uint64_t num_nodes;
};

// --- 1. keygen: compute_point_keys_and_sorted_indices (phase_1.cuh:62-95) -------
// One block per radix-sort tile (4096 points). The block also counts the sort's
// first digit (key bits 0-7) into per-wave LDS histograms, so that the sort skips
// its first counting pass. Full tiles of 16-byte aligned inputs are read with
// 16-byte loads and the keys written as uint4 (thread t owns points 4t..4t+3 of
// each quarter of the tile). Points are either two coordinate arrays (the column
// API) or interleaved x, y pairs (vec_2d, the header-only API's layout).
template <bool XY>
__global__ void __launch_bounds__(256)
  keygen_kernel(double const* __restrict__ xs,  // XY: interleaved x, y pairs
                double const* __restrict__ ys,  // XY: unused
                uint64_t n,
                double min_x,
                double min_y,
                double max_x,
                double max_y,
                double scale,
                uint32_t outside_key,
                bool vec,
                uint32_t* __restrict__ keys,
                uint32_t* __restrict__ counts)
{
  constexpr int NW = 256 / kWave;
  __shared__ uint32_t h[NW][kRadix];
  int const w = threadIdx.x / kWave;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) h[ww][threadIdx.x] = 0;
  __syncthreads();
  uint64_t const lo = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  if (vec && lo + kSortTile <= n) {
    constexpr int Q = kSortItems / 4;
    double px[kSortItems], py[kSortItems];
    if (XY) {
      double2 const* p2 = reinterpret_cast<double2 const*>(xs) + lo;
#pragma unroll
      for (int j = 0; j < Q; ++j) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          double2 const v   = p2[(j * 256 + threadIdx.x) * 4 + c];
          px[4 * j + c]     = v.x;
          py[4 * j + c]     = v.y;
        }
      }
    } else {
      double2 const* x2 = reinterpret_cast<double2 const*>(xs + lo);
      double2 const* y2 = reinterpret_cast<double2 const*>(ys + lo);
#pragma unroll
      for (int j = 0; j < Q; ++j) {
        int const e = (j * 256 + threadIdx.x) * 2;  // double2 index of point 4 * (j * 256 + t)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          double2 const xv  = x2[e + c];
          double2 const yv  = y2[e + c];
          px[4 * j + 2 * c] = xv.x;
          px[4 * j + 2 * c + 1] = xv.y;
          py[4 * j + 2 * c] = yv.x;
          py[4 * j + 2 * c + 1] = yv.y;
        }
      }
    }
    uint4* k4 = reinterpret_cast<uint4*>(keys + lo);
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      uint32_t k[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        k[c] = point_key(px[4 * j + c], py[4 * j + c], min_x, min_y, max_x, max_y, scale,
                         outside_key);
        wave_count_digit(h[w], k[c] & 0xFFu);
      }
      k4[j * 256 + threadIdx.x] = make_uint4(k[0], k[1], k[2], k[3]);
    }
  } else {
    double x[kSortItems], y[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      uint64_t const i = lo + static_cast<uint64_t>(j) * 256 + threadIdx.x;
      x[j]             = i < n ? (XY ? xs[2 * i] : xs[i]) : 0.0;
      y[j]             = i < n ? (XY ? xs[2 * i + 1] : ys[i]) : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      uint64_t const i   = lo + static_cast<uint64_t>(j) * 256 + threadIdx.x;
      uint32_t const key = point_key(x[j], y[j], min_x, min_y, max_x, max_y, scale, outside_key);
      if (i < n) keys[i] = key;
      wave_count_digit(h[w], i < n ? (key & 0xFFu) : static_cast<uint32_t>(kRadix));
    }
  }
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) s += h[ww][threadIdx.x];
  counts[static_cast<uint64_t>(blockIdx.x) * kRadix + threadIdx.x] = s;
}

// --- 3. level 0 and the root-level quirk ----------------------------------------
// One wave. Lane v tests the level-0 prefix v (key >> shift0 == v, v < 64).
__global__ void __launch_bounds__(64)
  level0_kernel(uint32_t const* __restrict__ K,
                uint32_t n,
                int d,
                uint32_t* __restrict__ nkey,
                uint32_t* __restrict__ nstart,
                uint32_t* __restrict__ ncount,
                uint8_t* __restrict__ nlevel,
                TreeMeta* __restrict__ meta)
{
  int const v        = lane_id();
  int const shift0   = d >= 1 ? 2 * (d - 1) : 0;
  uint32_t const lo  = lower_bound_u32(K, 0, n, static_cast<uint64_t>(v) << shift0);
  uint32_t const hi  = lower_bound_u32(K, lo, n, static_cast<uint64_t>(v + 1) << shift0);
  bool const present = hi > lo;
  uint64_t const m   = ballot(present);
  if (present) {
    uint32_t const slot = static_cast<uint32_t>(__popcll(m & lanemask_lt()));
    nkey[slot]          = static_cast<uint32_t>(v);
    nstart[slot]        = lo;
    ncount[slot]        = hi - lo;
    nlevel[slot]        = 0;
  }
  // Root level: distinct key >> 2d (d >= 1); at most 16 values.
  bool root_present = false;
  if (d >= 1 && v < 16) {
    uint32_t const rlo = lower_bound_u32(K, 0, n, static_cast<uint64_t>(v) << (2 * d));
    uint32_t const rhi = lower_bound_u32(K, rlo, n, static_cast<uint64_t>(v + 1) << (2 * d));
    root_present       = rhi > rlo;
  }
  uint64_t const rm = ballot(root_present);
  if (v == 0) {
    for (int l = 0; l <= kMaxLevels; ++l) meta->sizes[l] = 0;
    meta->sizes[0]     = static_cast<uint64_t>(__popcll(m));
    meta->overflow     = 0;
    int const R        = __popcll(rm);
    uint32_t const E   = (d >= 1 && R > 1) ? static_cast<uint32_t>(R - 1) : 0u;
    meta->num_marked   = E;
    // The first E distinct bottom-level keys (the extra "parents").
    uint32_t pos = 0;
    for (uint32_t e = 0; e < E && pos < n; ++e) {
      uint32_t const k      = K[pos];
      meta->marked_key[e]   = k;
      meta->marked_internal[e] = 0;
      meta->marked_start[e] = 0;
      meta->marked_count[e] = 0;
      pos                   = lower_bound_u32(K, pos, n, static_cast<uint64_t>(k) + 1);
    }
  }
}

// After all levels: total node count and the kept extra "parents".
__global__ void __launch_bounds__(64)
  post_levels_kernel(uint32_t const* __restrict__ nkey,
                     uint32_t const* __restrict__ nstart,
                     uint32_t const* __restrict__ ncount,
                     int d,
                     uint32_t max_size,
                     uint64_t cap,
                     TreeMeta* __restrict__ meta)
{
  if (threadIdx.x != 0) return;
  uint64_t total = 0, bottom_begin = 0;
  int const levels = d >= 1 ? d : 1;
  for (int l = 0; l < levels; ++l) {
    if (l == levels - 1) bottom_begin = total;
    total += meta->sizes[l];
  }
  meta->num_nodes = total;
  if (total > cap) {
    meta->overflow = 1;
    return;
  }
  uint32_t const lo0 = static_cast<uint32_t>(bottom_begin);
  uint32_t const hi0 = static_cast<uint32_t>(total);
  for (uint32_t e = 0; e < meta->num_marked; ++e) {
    uint32_t const mk = meta->marked_key[e];
    uint32_t const at = lower_bound_u32(nkey, lo0, hi0, mk);
    if (at < hi0 && nkey[at] == mk && ncount[at] > max_size) {
      meta->marked_internal[e] = 1;
      meta->marked_start[e]    = nstart[at];
      meta->marked_count[e]    = ncount[at];
    }
  }
}

// --- 5. finalize: the five output columns (detail/point_quadtree.cuh:88-133) ----
__global__ void __launch_bounds__(256)
  finalize_kernel(uint32_t const* __restrict__ nkey,
                  uint32_t const* __restrict__ nstart,
                  uint32_t const* __restrict__ ncount,
                  uint8_t const* __restrict__ nlevel,
                  uint32_t const* __restrict__ cfirst,
                  uint32_t const* __restrict__ ccount,
                  TreeMeta const* __restrict__ meta,
                  int d,
                  uint32_t max_size,
                  uint64_t m,
                  uint32_t* __restrict__ out_key,
                  uint8_t* __restrict__ out_level,
                  uint8_t* __restrict__ out_internal,
                  uint32_t* __restrict__ out_length,
                  uint32_t* __restrict__ out_offset)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    [&] {
      uint32_t const key   = nkey[i];
      int const level      = nlevel[i];
      uint32_t const count = ncount[i];
      uint32_t const start = nstart[i];
      bool internal        = level <= d - 2 && count > max_size;
      uint32_t length      = internal ? ccount[i] : count;
      uint32_t offset      = internal ? cfirst[i] : start;
      uint32_t const E     = meta->num_marked;
      if (E > 0 && !internal) {
        bool marked = false;
        uint32_t skipped = 0;
        for (uint32_t e = 0; e < E; ++e) {
          if (!meta->marked_internal[e]) continue;
          if (level == d - 1 && meta->marked_key[e] == key) marked = true;
          if (meta->marked_start[e] < start) skipped += meta->marked_count[e];
        }
        if (marked) {
          internal = true;
          length   = 0;
          offset   = static_cast<uint32_t>(m);
        } else {
          offset = start - skipped;
        }
      }
      out_key[i]      = key;
      out_level[i]    = static_cast<uint8_t>(level);
      out_internal[i] = internal ? 1 : 0;
      out_length[i]   = length;
      out_offset[i]   = offset;
    }();
}

}  // namespace

// Host orchestration. Returns the number of nodes; outputs via `oa`.
static int64_t quadtree_on_points_impl(double const* x,
                                       double const* y,
                                       bool interleaved,  // x holds x, y pairs; y unused
                                       int64_t n,
                                       double x_min,
                                       double x_max,
                                       double y_min,
                                       double y_max,
                                       double scale,
                                       int8_t max_depth,
                                       int32_t max_size,
                                       OutputAllocator& oa,
                                       hipStream_t stream,
                                       uint32_t** point_indices,
                                       uint32_t** key,
                                       uint8_t** level,
                                       uint8_t** is_internal,
                                       uint32_t** length,
                                       uint32_t** offset)
{
  QPIP_EXPECTS(n < (int64_t{1} << 32), "number of points must be less than 2^32");
  // detail/point_quadtree.cuh:259-268
  double const min_x = std::min(x_min, x_max), max_x = std::max(x_min, x_max);
  double const min_y = std::min(y_min, y_max), max_y = std::max(y_min, y_max);
  max_size           = std::max(1, max_size);
  max_depth          = std::max(int8_t{0}, std::min(int8_t{15}, max_depth));
  int const d        = max_depth;
  scale = std::max(scale, std::max(max_x - min_x, max_y - min_y) / ((1 << d) + 2));

  // Cell coordinates are below 2^d + 3, so keys have at most 2*bits(2^d + 2) bits.
  int cbits = 0;
  while ((uint64_t{1} << cbits) <= (uint64_t{1} << d) + 2) ++cbits;
  int const key_bits = std::min(32, 2 * cbits);
  uint32_t const outside_key = static_cast<uint32_t>((1u << (2 * d)) - 1u);

  uint64_t const un = static_cast<uint64_t>(n);
  uint64_t const tiles = sort_tiles(un);
  // point_indices is written by the sort's last pass straight into the output
  uint32_t* pi = oa.make<uint32_t>(un);
  dbuf<uint32_t> keys_a(un), keys_b(un), vals_tmp(un);
  dbuf<uint32_t> counts0(static_cast<size_t>(tiles) * kRadix);
  {
    KernelTimer kt("keygen", stream);
    bool const vec = reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
                     (interleaved || reinterpret_cast<uintptr_t>(y) % 16 == 0) &&
                     reinterpret_cast<uintptr_t>(keys_a.get()) % 16 == 0;
    if (interleaved)
      keygen_kernel<true><<<static_cast<unsigned>(tiles), 256, 0, stream>>>(
        x, nullptr, un, min_x, min_y, max_x, max_y, scale, outside_key, vec, keys_a.get(),
        counts0.get());
    else
      keygen_kernel<false><<<static_cast<unsigned>(tiles), 256, 0, stream>>>(
        x, y, un, min_x, min_y, max_x, max_y, scale, outside_key, vec, keys_a.get(),
        counts0.get());
    QPIP_LAUNCH_CHECK();
  }
  SortResult sorted = radix_sort_pairs(keys_a.get(), keys_b.get(), nullptr, pi, vals_tmp.get(),
                                       un, key_bits, stream, counts0.get(), "point_sort");
  vals_tmp.release();

  // Node capacity: level 0 holds <= 64 nodes; level l >= 1 holds <= 4 children
  // of each internal node of level l-1 (internal nodes of a level are disjoint
  // and hold > max_size points each), <= n, and <= 64 * 4^l.
  uint64_t const internal_bound = un / (static_cast<uint64_t>(max_size) + 1);
  uint64_t cap                  = 64;
  for (int l = 1; l < std::max(d, 1); ++l) {
    uint64_t c = std::min<uint64_t>(un, 4 * internal_bound);
    if (2 * l + 6 < 62) c = std::min<uint64_t>(c, uint64_t{64} << (2 * l));
    cap += c;
  }
  dbuf<uint32_t> nkey(cap), nstart(cap), ncount(cap), cfirst(cap), ccount(cap);
  dbuf<uint8_t> nlevel(cap);
  dbuf<TreeMeta> meta(1);
  dbuf<uint4> split(cap);
  dbuf<uint32_t> partials(kLevelBlocks);

  {
  KernelTimer kt_levels("tree_levels", stream);
  level0_kernel<<<1, 64, 0, stream>>>(sorted.keys, static_cast<uint32_t>(un), d, nkey.get(),
                                      nstart.get(), ncount.get(), nlevel.get(), meta.get());
  QPIP_LAUNCH_CHECK();
  // Levels 0..d-2 may have children (level d-1 is the bottom).
  for (int l = 0; l + 2 <= d; ++l) {
    LevelArgs la{sorted.keys, nkey.get(), nstart.get(), ncount.get(), nlevel.get(),
                 cfirst.get(), ccount.get(), split.get(), partials.get(), meta.get()->sizes,
                 &meta.get()->overflow, cap, l};
    UniformLevels rule{d, static_cast<uint32_t>(max_size)};
    level_split_kernel<<<kLevelBlocks, 256, 0, stream>>>(la, rule);
    QPIP_LAUNCH_CHECK();
    level_emit_kernel<<<kLevelBlocks, 256, 0, stream>>>(la, rule);
    QPIP_LAUNCH_CHECK();
  }
  post_levels_kernel<<<1, 64, 0, stream>>>(nkey.get(), nstart.get(), ncount.get(), d,
                                           static_cast<uint32_t>(max_size), cap, meta.get());
  QPIP_LAUNCH_CHECK();
  }

  TreeMeta h_meta;
  QPIP_HIP(hipMemcpyAsync(&h_meta, meta.get(), sizeof(TreeMeta), hipMemcpyDeviceToHost, stream));
  QPIP_HIP(hipStreamSynchronize(stream));
  if (h_meta.overflow) throw device_error("quadtree node capacity exceeded (internal error)");
  uint64_t const m = h_meta.num_nodes;

  uint32_t* o_key      = oa.make<uint32_t>(m);
  uint8_t* o_level     = oa.make<uint8_t>(m);
  uint8_t* o_internal  = oa.make<uint8_t>(m);
  uint32_t* o_length   = oa.make<uint32_t>(m);
  uint32_t* o_offset   = oa.make<uint32_t>(m);
  if (m > 0) {
    KernelTimer kt("tree_finalize", stream);
    finalize_kernel<<<grid_for(m, 256), 256, 0, stream>>>(
      nkey.get(), nstart.get(), ncount.get(), nlevel.get(), cfirst.get(), ccount.get(),
      meta.get(), d, static_cast<uint32_t>(max_size), m, o_key, o_level, o_internal, o_length,
      o_offset);
    QPIP_LAUNCH_CHECK();
  }
  // no wait here: the outputs are ready for work queued on `stream` after this call
  *point_indices = pi;
  *key           = o_key;
  *level         = o_level;
  *is_internal   = o_internal;
  *length        = o_length;
  *offset        = o_offset;
  return static_cast<int64_t>(m);
}

}  // namespace qpip

namespace {

int quadtree_on_points_entry(const double* x,
                             const double* y,
                             bool interleaved,
                             int64_t num_points,
                             int64_t num_y,
                             double x_min,
                             double x_max,
                             double y_min,
                             double y_max,
                             double scale,
                             int8_t max_depth,
                             int32_t max_size,
                             const qpip_resource* mr,
                             void* stream,
                             uint32_t** point_indices,
                             uint32_t** key,
                             uint8_t** level,
                             uint8_t** is_internal_node,
                             uint32_t** length,
                             uint32_t** offset,
                             int64_t* num_nodes)
{
  using namespace qpip;
  try {
    QPIP_EXPECTS(point_indices && key && level && is_internal_node && length && offset && num_nodes,
                 "output pointers must not be null");
    *point_indices = nullptr;
    *key = nullptr, *level = nullptr, *is_internal_node = nullptr, *length = nullptr,
    *offset = nullptr, *num_nodes = 0;
    // cpp/src/indexing/point_quadtree.cu:166
    QPIP_EXPECTS(num_points == num_y, "x and y columns must have the same length");
    if (num_points <= 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(x && (interleaved || y), "x and y must not be null");
    StreamScope scope(static_cast<hipStream_t>(stream));
    OutputAllocator oa(mr, static_cast<hipStream_t>(stream));
    *num_nodes = quadtree_on_points_impl(x, y, interleaved, num_points, x_min, x_max, y_min,
                                         y_max, scale, max_depth, max_size, oa,
                                         static_cast<hipStream_t>(stream), point_indices, key,
                                         level, is_internal_node, length, offset);
    oa.keep();
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}

}  // namespace

extern "C" int qpip_quadtree_on_points(const double* x,
                                       const double* y,
                                       int64_t num_x,
                                       int64_t num_y,
                                       double x_min,
                                       double x_max,
                                       double y_min,
                                       double y_max,
                                       double scale,
                                       int8_t max_depth,
                                       int32_t max_size,
                                       const qpip_resource* mr,
                                       void* stream,
                                       uint32_t** point_indices,
                                       uint32_t** key,
                                       uint8_t** level,
                                       uint8_t** is_internal_node,
                                       uint32_t** length,
                                       uint32_t** offset,
                                       int64_t* num_nodes)
{
  return quadtree_on_points_entry(x, y, false, num_x, num_y, x_min, x_max, y_min, y_max, scale,
                                  max_depth, max_size, mr, stream, point_indices, key, level,
                                  is_internal_node, length, offset, num_nodes);
}

extern "C" int qpip_quadtree_on_points_xy(const double* points_xy,
                                          int64_t num_points,
                                          double x_min,
                                          double x_max,
                                          double y_min,
                                          double y_max,
                                          double scale,
                                          int8_t max_depth,
                                          int32_t max_size,
                                          const qpip_resource* mr,
                                          void* stream,
                                          uint32_t** point_indices,
                                          uint32_t** key,
                                          uint8_t** level,
                                          uint8_t** is_internal_node,
                                          uint32_t** length,
                                          uint32_t** offset,
                                          int64_t* num_nodes)
{
  return quadtree_on_points_entry(points_xy, nullptr, true, num_points, num_points, x_min, x_max, y_min,
                                  y_max, scale, max_depth, max_size, mr, stream, point_indices,
                                  key, level, is_internal_node, length, offset, num_nodes);
}
