// Stable radix sort and stream compaction exposed through the C ABI, for the
// library's own callers above it (the learned locator and the serving loop order
// and compact their results with these instead of a vendor primitive).
//
// qpip_radix_sort_pairs is the stable LSD sort of the quadtree path (qpip_sort.hip),
// the counterpart of the reference's thrust::stable_sort_by_key on uint32 keys
// (phase_1.cuh:92, quadtree_bbox_filtering.cuh:161-165).
// qpip_compact_nonnegative is the count -> scan -> emit sequence (qpip_scan.hpp),
// the counterpart of a stable thrust::copy_if (quadtree_point_in_polygon.cuh:198-207).
#include <algorithm>

#include "qpip_scan.hpp"

namespace qpip {

namespace {

__global__ void copy_u32_kernel(uint32_t const* __restrict__ src, uint32_t* __restrict__ dst,
                                uint64_t n)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// Stable compaction of the entries with a non-negative value: their indices and
// values, in index order.
struct NonNegativeOp {
  int32_t const* values;
  uint32_t* indices_out;
  int32_t* values_out;
  __device__ uint32_t count(uint64_t i) const { return values[i] >= 0 ? 1u : 0u; }
  __device__ void emit(uint64_t i, uint32_t prefix) const
  {
    int32_t const v = values[i];
    if (v < 0) return;
    indices_out[prefix] = static_cast<uint32_t>(i);
    if (values_out) values_out[prefix] = v;
  }
};

}  // namespace

}  // namespace qpip

extern "C" int qpip_radix_sort_pairs(uint32_t* keys,
                                     uint32_t* values,
                                     int64_t n,
                                     int32_t key_bits,
                                     void* stream_)
{
  using namespace qpip;
  return guarded([&] {
    QPIP_EXPECTS(n >= 0, "number of pairs must not be negative");
    QPIP_EXPECTS(key_bits >= 1 && key_bits <= 32, "key bits must be in [1, 32]");
    if (n == 0) return;
    QPIP_EXPECTS(keys && values, "keys and values must not be null");
    QPIP_EXPECTS(n < (int64_t{1} << 32), "at most 2^32 - 1 pairs");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    StreamScope scope(stream);
    uint64_t const un = static_cast<uint64_t>(n);
    dbuf<uint32_t> kb(un), vin(un), vtmp(un), vout(un);
    unsigned const g = grid_for(un, 256, 4096);
    copy_u32_kernel<<<g, 256, 0, stream>>>(values, vin.get(), un);
    QPIP_LAUNCH_CHECK();
    SortResult r = radix_sort_pairs(keys, kb.get(), vin.get(), vout.get(), vtmp.get(), un,
                                    key_bits, stream, nullptr, "sort_pairs");
    if (r.keys != keys) {
      copy_u32_kernel<<<g, 256, 0, stream>>>(r.keys, keys, un);
      QPIP_LAUNCH_CHECK();
    }
    copy_u32_kernel<<<g, 256, 0, stream>>>(r.values, values, un);
    QPIP_LAUNCH_CHECK();
  });
}

extern "C" int qpip_compact_nonnegative(const int32_t* values,
                                        int64_t n,
                                        uint32_t* indices_out,
                                        int32_t* values_out,
                                        int64_t* count,
                                        void* stream_)
{
  using namespace qpip;
  return guarded([&] {
    QPIP_EXPECTS(count != nullptr, "output pointers must not be null");
    *count = 0;
    QPIP_EXPECTS(n >= 0, "number of values must not be negative");
    if (n == 0) return;
    QPIP_EXPECTS(values && indices_out, "values and indices_out must not be null");
    QPIP_EXPECTS(n < (int64_t{1} << 32), "at most 2^32 - 1 values");
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    StreamScope scope(stream);
    uint64_t const un = static_cast<uint64_t>(n);
    dbuf<uint32_t> partials(kScanBlocks + 1), total(1);
    NonNegativeOp op{values, indices_out, values_out};
    count_scan_emit<NonNegativeOp, uint32_t, uint32_t>(op, nullptr, un, partials.get(),
                                                       total.get(), stream, scan_blocks_for(un),
                                                       "compact");
    uint32_t h = 0;
    QPIP_HIP(hipMemcpyAsync(&h, total.get(), sizeof(h), hipMemcpyDeviceToHost, stream));
    QPIP_HIP(hipStreamSynchronize(stream));
    *count = static_cast<int64_t>(h);
  });
}
