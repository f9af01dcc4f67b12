// join_quadtree_and_bounding_boxes on gfx950.
//
// Reference: cpp/src/join/quadtree_bbox_filtering.cu:90-126 (column API, checks)
// and detail/join/quadtree_bbox_filtering.cuh:37-188 with intersection.cuh:79-136
// and traversal.cuh:63-145: a breadth-first traversal that starts from every
// (level-0 node, bbox) pair (node = i % T, bbox = i / T), classifies each pair
// as leaf / quad / none, emits leaf pairs in order, expands quad pairs to their
// children in order, and finally stable-sorts the leaf pairs by the leaf's first
// point position.
//
// Here the same traversal runs entirely on the device, one count/scan/emit per
// level with the frontier length kept in device memory (the reference syncs the
// host on every level for sizes, traversal.cuh:222). Each frontier element counts
// its leaf and quad children in one packed 64-bit value, so a single scan gives
// both output positions and the emission order is the reference's. The stable
// sort is the library's LSD radix sort over the bits of the largest leaf offset.
// Frontier/output capacities are preallocated; on overflow the traversal is
// rerun with the sizes the device recorded.
#include <algorithm>

#include "qpip_join.hpp"

namespace qpip {

namespace {

// Level 0: the virtual frontier i -> (i % T, i / T), each pair classified once.
struct TopOp {
  TreeView t;
  BoxView bx;
  Params p;
  JoinMeta* meta;
  uint32_t* out_node;
  uint32_t* out_bbox;
  uint32_t* fr_node;
  uint32_t* fr_bbox;
  uint64_t cap_out, cap_fr;

  __device__ uint64_t count(uint64_t i) const
  {
    uint32_t const T = meta->num_top;
    int const c      = classify(t, bx, p, static_cast<uint32_t>(i % T), static_cast<uint32_t>(i / T));
    return c == 0 ? 1ull : (c == 1 ? (1ull << 32) : 0ull);
  }
  __device__ void emit(uint64_t i, uint64_t prefix) const
  {
    uint32_t const T    = meta->num_top;
    uint32_t const node = static_cast<uint32_t>(i % T);
    uint32_t const bbox = static_cast<uint32_t>(i / T);
    int const c         = classify(t, bx, p, node, bbox);
    if (c == 0) {
      uint64_t const o = prefix & 0xFFFFFFFFull;
      if (o < cap_out) {
        out_node[o] = node;
        out_bbox[o] = bbox;
      } else {
        atomicOr(&meta->overflow, 1u);
      }
    } else if (c == 1) {
      uint64_t const f = prefix >> 32;
      if (f < cap_fr) {
        fr_node[f] = node;
        fr_bbox[f] = bbox;
      } else {
        atomicOr(&meta->overflow, 1u);
      }
    }
  }
};

// T = count(level == 0) (quadtree_bbox_filtering.cuh:53-56) and the largest
// leaf offset (for the sort's bit count).
struct TopCountOp {
  uint8_t const* level;
  __device__ uint32_t count(uint64_t i) const { return level[i] == 0 ? 1u : 0u; }
  __device__ void emit(uint64_t, uint32_t) const {}
};

__global__ void __launch_bounds__(256)
  leaf_offset_max_kernel(uint8_t const* __restrict__ internal, uint32_t const* __restrict__ offset,
                         uint64_t m, JoinMeta* __restrict__ meta)
{
  uint32_t mx = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    if (!internal[i]) mx = max(mx, offset[i]);
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) mx = max(mx, static_cast<uint32_t>(__shfl_xor(mx, o, kWave)));
  if (lane_id() == 0) atomicMax(&meta->max_leaf_offset, mx);
}

__global__ void setup_kernel(JoinMeta* meta, uint32_t const* num_top, uint64_t b)
{
  meta->num_top   = *num_top;
  meta->top_pairs = static_cast<uint64_t>(*num_top) * b;
}

__global__ void __launch_bounds__(256)
  gather_keys_kernel(uint32_t const* __restrict__ out_node, uint32_t const* __restrict__ offset,
                     uint64_t k, uint32_t* __restrict__ keys)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    keys[i] = offset[out_node[i]];
}

__global__ void __launch_bounds__(256)
  permute_pairs_kernel(uint32_t const* __restrict__ perm, uint32_t const* __restrict__ out_node,
                       uint32_t const* __restrict__ out_bbox, uint64_t k,
                       uint32_t* __restrict__ bbox_offset, uint32_t* __restrict__ quad_offset)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t const j = perm[i];
    bbox_offset[i]   = out_bbox[j];
    quad_offset[i]   = out_node[j];
  }
}

}  // namespace

}  // namespace qpip

extern "C" int qpip_join_quadtree_and_bounding_boxes(const uint32_t* key,
                                                     const uint8_t* level,
                                                     const uint8_t* is_internal_node,
                                                     const uint32_t* length,
                                                     const uint32_t* offset,
                                                     int64_t num_nodes,
                                                     const double* bbox_minx,
                                                     const double* bbox_miny,
                                                     const double* bbox_maxx,
                                                     const double* bbox_maxy,
                                                     int64_t num_bboxes,
                                                     double x_min,
                                                     double x_max,
                                                     double y_min,
                                                     double y_max,
                                                     double scale,
                                                     int8_t max_depth,
                                                     const qpip_resource* mr,
                                                     void* stream_,
                                                     uint32_t** bbox_offset,
                                                     uint32_t** quad_offset,
                                                     int64_t* num_pairs)
{
  using namespace qpip;
  try {
    QPIP_EXPECTS(bbox_offset && quad_offset && num_pairs, "output pointers must not be null");
    *bbox_offset = *quad_offset = nullptr;
    *num_pairs                  = 0;
    hipStream_t stream          = static_cast<hipStream_t>(stream_);
    // quadtree_bbox_filtering.cu:100-114
    QPIP_EXPECTS(scale > 0, "scale must be positive");
    QPIP_EXPECTS(x_min < x_max && y_min < y_max, "invalid bounding box (x_min, x_max, y_min, y_max)");
    QPIP_EXPECTS(max_depth > 0 && max_depth < 16, "maximum depth must be positive and less than 16");
    if (num_nodes <= 0 || num_bboxes <= 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(num_nodes < (int64_t{1} << 32) && num_bboxes < (int64_t{1} << 32),
                 "quadtree and bounding box tables must have fewer than 2^32 rows");

    StreamScope scope(stream);
    TreeView tv{key, level, is_internal_node, length, offset, static_cast<uint64_t>(num_nodes)};
    BoxView bv{bbox_minx, bbox_miny, bbox_maxx, bbox_maxy, static_cast<uint64_t>(num_bboxes)};
    Params prm{x_min, y_min, scale, static_cast<int>(max_depth)};
    uint64_t const M = static_cast<uint64_t>(num_nodes);
    uint64_t const B = static_cast<uint64_t>(num_bboxes);

    dbuf<JoinMeta> meta(1);
    dbuf<uint64_t> partials64(kScanBlocks + 1);
    dbuf<uint32_t> partials32(kScanBlocks + 1);
    dbuf<uint32_t> num_top(1);

    uint64_t cap_out = std::max<uint64_t>(uint64_t{1} << 20, 8 * B);
    uint64_t cap_fr  = std::max<uint64_t>(uint64_t{1} << 20, 8 * B);
    JoinMeta h_meta{};
    dbuf<uint32_t> out_node, out_bbox, fa_node, fa_bbox, fb_node, fb_bbox;
    for (int attempt = 0;; ++attempt) {
      out_node.reset(cap_out);
      out_bbox.reset(cap_out);
      fa_node.reset(cap_fr);
      fa_bbox.reset(cap_fr);
      fb_node.reset(cap_fr);
      fb_bbox.reset(cap_fr);
      QPIP_HIP(hipMemsetAsync(meta.get(), 0, sizeof(JoinMeta), stream));
      count_scan_emit<TopCountOp, uint32_t, uint32_t>(TopCountOp{level}, nullptr, M,
                                                      partials32.get(), num_top.get(), stream,
                                                      scan_blocks_for(M), "join_setup");
      setup_kernel<<<1, 1, 0, stream>>>(meta.get(), num_top.get(), B);
      QPIP_LAUNCH_CHECK();
      leaf_offset_max_kernel<<<grid_for(M, 256, 1024), 256, 0, stream>>>(is_internal_node, offset,
                                                                         M, meta.get());
      QPIP_LAUNCH_CHECK();
      JoinMeta* dm = meta.get();
      TopOp top{tv, bv, prm, dm, out_node.get(), out_bbox.get(), fa_node.get(), fa_bbox.get(),
                cap_out, cap_fr};
      count_scan_emit<TopOp, uint64_t, SplitTotal>(top, &dm->top_pairs, 0, partials64.get(),
                                                   &dm->level[0], stream, kScanBlocks,
                                                   "join_level");
      uint32_t* cur_node = fa_node.get();
      uint32_t* cur_bbox = fa_bbox.get();
      uint32_t* nxt_node = fb_node.get();
      uint32_t* nxt_bbox = fb_bbox.get();
      // quadtree_bbox_filtering.cuh:116-163: levels 1 .. max_depth-1
      for (int lvl = 1; lvl < max_depth; ++lvl) {
        DescendOp<Params> op{tv, bv, prm, dm, cur_node, cur_bbox, out_node.get(), out_bbox.get(),
                             nxt_node, nxt_bbox, cap_out, cap_fr, lvl};
        count_scan_emit<DescendOp<Params>, uint64_t, SplitTotal>(op, &dm->level[lvl - 1].quad, 0,
                                                         partials64.get(), &dm->level[lvl], stream,
                                                         kScanBlocks, "join_level");
        std::swap(cur_node, nxt_node);
        std::swap(cur_bbox, nxt_bbox);
      }
      QPIP_HIP(hipMemcpyAsync(&h_meta, dm, sizeof(JoinMeta), hipMemcpyDeviceToHost, stream));
      QPIP_HIP(hipStreamSynchronize(stream));
      if (h_meta.bad_index)
        throw logic_error("cuSpatial failure: quadtree child range exceeds the number of nodes");
      if (!h_meta.overflow) break;
      // grow to what the device saw (plus headroom) and rerun
      uint64_t tot_out = 0, max_fr = h_meta.top_pairs;
      for (int l = 0; l < max_depth; ++l) {
        tot_out += h_meta.level[l].leaf;
        max_fr = std::max<uint64_t>(max_fr, h_meta.level[l].quad);
      }
      cap_out = std::max<uint64_t>(cap_out * 2, tot_out + tot_out / 4 + 1024);
      cap_fr  = std::max<uint64_t>(cap_fr * 2, max_fr + max_fr / 4 + 1024);
      if (attempt > 8) throw device_error("join capacity growth did not converge");
    }
    uint64_t K = 0;
    for (int l = 0; l < max_depth; ++l) K += h_meta.level[l].leaf;
    if (K == 0) return QPIP_SUCCESS;
    QPIP_EXPECTS(K < (uint64_t{1} << 32), "join produced more than 2^32 pairs");

    // stable_sort_by_key(offset[leaf], (bbox, leaf)) (quadtree_bbox_filtering.cuh:152-166)
    dbuf<uint32_t> ka(K), kb(K), va(K), vb(K);
    gather_keys_kernel<<<grid_for(K, 256), 256, 0, stream>>>(out_node.get(), offset, K, ka.get());
    QPIP_LAUNCH_CHECK();
    int bits = 1;
    while (bits < 32 && (h_meta.max_leaf_offset >> bits) != 0) ++bits;
    SortResult s = radix_sort_pairs(ka.get(), kb.get(), nullptr, va.get(), vb.get(), K, bits,
                                    stream, nullptr, "join_sort");
    OutputAllocator oa(mr, stream);
    uint32_t* o_bbox = oa.make<uint32_t>(K);
    uint32_t* o_quad = oa.make<uint32_t>(K);
    permute_pairs_kernel<<<grid_for(K, 256), 256, 0, stream>>>(s.values, out_node.get(),
                                                               out_bbox.get(), K, o_bbox, o_quad);
    QPIP_LAUNCH_CHECK();
    oa.keep();
    *bbox_offset = o_bbox;
    *quad_offset = o_quad;
    *num_pairs   = static_cast<int64_t>(K);
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}
