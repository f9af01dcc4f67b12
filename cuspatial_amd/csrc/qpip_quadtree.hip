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
//   1. keygen (fused with pass 0's tile histogram of the radix sort);
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

#include "qpip_scan.hpp"
#include "qpip_sort.hpp"

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

// --- z-order (detail/utility/z_order.cuh:25-77): spread bit i to bit 2i. --------
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

// --- 1. keygen: compute_point_keys_and_sorted_indices (phase_1.cuh:62-95) -------
// Tile layout is the radix sort's (256 threads x 16 items, wave-striped) so the
// digit-0 histogram of each tile can be written here.
__global__ void __launch_bounds__(kSortThreads)
  keygen_kernel(double const* __restrict__ xs,
                double const* __restrict__ ys,
                uint64_t n,
                double min_x,
                double min_y,
                double max_x,
                double max_y,
                double scale,
                uint32_t outside_key,
                uint32_t digit_mask,
                uint32_t* __restrict__ keys,
                uint32_t* __restrict__ counts,
                uint64_t tiles)
{
  __shared__ uint32_t hist[kRadix];
  for (int i = threadIdx.x; i < kRadix; i += kSortThreads) hist[i] = 0;
  __syncthreads();
  uint64_t const tile = blockIdx.x;
  int const w         = threadIdx.x / kWave;
  uint64_t const base = tile * kSortTile + static_cast<uint64_t>(w) * (kSortItems * kWave);
#pragma unroll 4
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t const idx = base + static_cast<uint64_t>(j) * kWave + lane_id();
    bool const valid   = idx < n;
    uint32_t key       = 0;
    if (valid) {
      double const x = xs[idx];
      double const y = ys[idx];
      if (x < min_x || x > max_x || y < min_y || y > max_y) {
        key = outside_key;
      } else {
        uint32_t const cx = to_u16((x - min_x) / scale);
        uint32_t const cy = to_u16((y - min_y) / scale);
        key               = (dilate_2(cy) << 1) | dilate_2(cx);
      }
      keys[idx] = key;
    }
    wave_hist_add(hist, key & digit_mask, valid);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) counts[d * tiles + tile] = hist[d];
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

// --- 4. level expansion (one level, count/scan/emit) ----------------------------
struct ExpandOp {
  uint32_t const* K;
  uint32_t* nkey;
  uint32_t* nstart;
  uint32_t* ncount;
  uint8_t* nlevel;
  uint32_t* cfirst;
  uint32_t* ccount;
  TreeMeta* meta;
  uint64_t cap;
  int level;
  int d;
  uint32_t max_size;

  __device__ __forceinline__ uint64_t begin() const
  {
    uint64_t b = 0;
    for (int l = 0; l < level; ++l) b += meta->sizes[l];
    return b;
  }

  // Child quadrant boundaries of node `node`: b[q]..b[q+1] is quadrant q.
  __device__ __forceinline__ uint32_t split(uint64_t node, uint32_t b[5]) const
  {
    if (node >= cap) return 0;  // only after an overflow (reported to the host)
    uint32_t const c = ncount[node];
    if (c <= max_size) return 0;
    uint32_t const k  = nkey[node];
    uint32_t const s  = nstart[node];
    uint32_t const e  = s + c;
    int const sh      = 2 * (d - 2 - level);
    b[0]              = s;
    b[4]              = e;
    uint32_t nonempty = 0;
#pragma unroll
    for (int q = 1; q < 4; ++q)
      b[q] = lower_bound_u32(K, b[q - 1], e, (static_cast<uint64_t>(k) * 4 + q) << sh);
#pragma unroll
    for (int q = 0; q < 4; ++q) nonempty += b[q + 1] > b[q];
    return nonempty;
  }

  __device__ uint32_t count(uint64_t i) const
  {
    uint32_t b[5];
    return split(begin() + i, b);
  }

  __device__ void emit(uint64_t i, uint32_t prefix) const
  {
    uint64_t const b0   = begin();
    uint64_t const node = b0 + i;
    if (node >= cap) return;
    uint32_t b[5];
    uint32_t const nc   = split(node, b);
    uint64_t const next = b0 + meta->sizes[level] + prefix;  // first child slot
    cfirst[node]        = static_cast<uint32_t>(next);
    ccount[node]        = nc;
    if (nc == 0) return;
    if (next + nc > cap) {
      atomicOr(&meta->overflow, 1u);
      return;
    }
    uint32_t const k = nkey[node];
    uint64_t slot    = next;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (b[q + 1] > b[q]) {
        nkey[slot]   = k * 4 + q;
        nstart[slot] = b[q];
        ncount[slot] = b[q + 1] - b[q];
        nlevel[slot] = static_cast<uint8_t>(level + 1);
        ++slot;
      }
    }
  }
};

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
  uint64_t const i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
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
}

}  // namespace

// Host orchestration. Returns the number of nodes; outputs via `oa`.
static int64_t quadtree_on_points_impl(double const* x,
                                       double const* y,
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

  uint64_t const un    = static_cast<uint64_t>(n);
  uint64_t const tiles = sort_tiles(un);
  dbuf<uint32_t> keys_a(un), keys_b(un), vals_a(un), vals_b(un);
  dbuf<uint32_t> counts(static_cast<size_t>(kRadix) * tiles);
  int const nb0 = std::min(kRadixBits, key_bits);
  {
    KernelTimer kt("keygen", stream);
    keygen_kernel<<<static_cast<unsigned>(tiles), kSortThreads, 0, stream>>>(
    x, y, un, min_x, min_y, max_x, max_y, scale, outside_key, (1u << nb0) - 1u, keys_a.get(),
    counts.get(), tiles);
    QPIP_LAUNCH_CHECK();
  }
  SortResult sorted = radix_sort_pairs(keys_a.get(), keys_b.get(), nullptr, vals_a.get(),
                                       vals_b.get(), un, key_bits, stream, counts.get(),
                                       "point_sort");
  counts.release();

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
  dbuf<uint32_t> partials(kScanBlocks + 1);

  {
  KernelTimer kt_levels("tree_levels", stream);
  level0_kernel<<<1, 64, 0, stream>>>(sorted.keys, static_cast<uint32_t>(un), d, nkey.get(),
                                      nstart.get(), ncount.get(), nlevel.get(), meta.get());
  QPIP_LAUNCH_CHECK();
  // Levels 0..d-2 may have children (level d-1 is the bottom).
  for (int l = 0; l + 2 <= d; ++l) {
    ExpandOp op{sorted.keys, nkey.get(), nstart.get(), ncount.get(), nlevel.get(),
                cfirst.get(), ccount.get(), meta.get(), cap, l, d,
                static_cast<uint32_t>(max_size)};
    // sizes[l] is this level's length; sizes[l+1] receives the number of children
    uint64_t* sizes = meta.get()->sizes;
    count_scan_emit<ExpandOp, uint32_t, uint64_t>(op, sizes + l, 0, partials.get(), sizes + l + 1,
                                                  stream, kScanBlocks, "tree_level_scan");
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
  // point_indices: the sorted values buffer
  uint32_t* pi = nullptr;
  if (sorted.values == vals_a.get())
    pi = oa.adopt(vals_a, un);
  else
    pi = oa.adopt(vals_b, un);
  QPIP_HIP(hipStreamSynchronize(stream));
  *point_indices = pi;
  *key           = o_key;
  *level         = o_level;
  *is_internal   = o_internal;
  *length        = o_length;
  *offset        = o_offset;
  return static_cast<int64_t>(m);
}

}  // namespace qpip

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
  using namespace qpip;
  try {
    QPIP_EXPECTS(point_indices && key && level && is_internal_node && length && offset && num_nodes,
                 "output pointers must not be null");
    *point_indices = nullptr;
    *key = nullptr, *level = nullptr, *is_internal_node = nullptr, *length = nullptr,
    *offset = nullptr, *num_nodes = 0;
    // cpp/src/indexing/point_quadtree.cu:166-177
    QPIP_EXPECTS(num_x == num_y, "x and y columns must have the same length");
    if (num_x <= 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(x && y, "x and y must not be null");
    OutputAllocator oa(mr, static_cast<hipStream_t>(stream));
    *num_nodes = quadtree_on_points_impl(x, y, num_x, x_min, x_max, y_min, y_max, scale,
                                         max_depth, max_size, oa, static_cast<hipStream_t>(stream),
                                         point_indices, key, level, is_internal_node, length,
                                         offset);
    return QPIP_SUCCESS;
  } catch (logic_error const& e) {
    set_last_error(e.what());
    return QPIP_ERROR_LOGIC;
  } catch (oom_error const& e) {
    set_last_error(e.what());
    return QPIP_ERROR_OOM;
  } catch (std::exception const& e) {
    set_last_error(e.what());
    return QPIP_ERROR_DEVICE;
  }
}
