// polygon_bounding_boxes on gfx950.
//
// Reference: cpp/src/bounding_boxes/polygon_bounding_boxes.cu:132-159 (column
// API) and detail/bounding_boxes.cuh:62-90,139-184: a reduce_by_key of per-vertex
// boxes keyed by each vertex's polygon id, id(v) = upper_bound(first, v) over the
// first ring offset of every polygon (iterator_factory.cuh:168-182,415-426).
//
// Equivalent formulation used here: the ids split [0, num_vertices) into the
// intervals [first[i-1], first[i]) (with [0, first[0]) and [first[P], nv) at the
// ends); reduce_by_key emits one box per non-empty interval, in order. So:
//   1. count/scan/emit over the P + 2 intervals assigns each non-empty interval
//      its output row (rows >= P are dropped: the reference writes past the end);
//   2. one wave per output row reduces its vertices with DPP/shuffle min/max.
// For well-formed input (no empty polygon, offsets starting at 0) interval i+1
// is polygon i. Rows that receive no interval are NaN (uninitialised in the
// reference).
#include <cmath>

#include <algorithm>

#include "qpip_box.hpp"
#include "qpip_scan.hpp"

namespace qpip {

namespace {

struct IntervalOp {
  int32_t const* poly_offsets;
  int32_t const* ring_offsets;
  int64_t num_poly_offsets;  // P + 1
  int64_t num_ring_offsets;
  int64_t nv;
  int64_t num_polys;  // P
  int64_t* lo_out;
  int64_t* hi_out;
  uint32_t* bad;

  __device__ int64_t first(int64_t i) const
  {
    int64_t const po = poly_offsets[i];
    if (po < 0 || po >= num_ring_offsets) {
      atomicOr(bad, 1u);
      return 0;
    }
    return ring_offsets[po];
  }
  __device__ void range(uint64_t i, int64_t& lo, int64_t& hi) const
  {
    int64_t const P = num_poly_offsets - 1;
    lo              = i == 0 ? 0 : first(static_cast<int64_t>(i) - 1);
    hi              = static_cast<int64_t>(i) <= P ? first(static_cast<int64_t>(i)) : nv;
    if (lo < 0) lo = 0;
    if (hi > nv) hi = nv;
  }
  __device__ uint32_t count(uint64_t i) const
  {
    int64_t lo, hi;
    range(i, lo, hi);
    return hi > lo ? 1u : 0u;
  }
  __device__ void emit(uint64_t i, uint32_t row) const
  {
    int64_t lo, hi;
    range(i, lo, hi);
    if (hi > lo && static_cast<int64_t>(row) < num_polys) {
      lo_out[row] = lo;
      hi_out[row] = hi;
    }
  }
};

__global__ void __launch_bounds__(256)
  bbox_reduce_kernel(double const* __restrict__ x,
                     double const* __restrict__ y,
                     int64_t const* __restrict__ lo_in,
                     int64_t const* __restrict__ hi_in,
                     uint32_t const* __restrict__ num_runs,
                     int64_t num_polys,
                     double r,
                     double* __restrict__ minx,
                     double* __restrict__ miny,
                     double* __restrict__ maxx,
                     double* __restrict__ maxy,
                     int64_t base)
{
  int64_t const row =
    base + static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (row >= num_polys) return;
  double const nan = __builtin_nan("");
  if (row >= static_cast<int64_t>(*num_runs)) {
    if (lane_id() == 0) minx[row] = miny[row] = maxx[row] = maxy[row] = nan;
    return;
  }
  int64_t const lo = lo_in[row], hi = hi_in[row];
  double lx = nan, ly = nan, hx = nan, hy = nan;
  for (int64_t v = lo + lane_id(); v < hi; v += kWave) {
    double const vx = x[v], vy = y[v];
    lx              = box_min(lx, vx - r);
    ly              = box_min(ly, vy - r);
    hx              = box_max(hx, vx + r);
    hy              = box_max(hy, vy + r);
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    lx = box_min(lx, __shfl_xor(lx, o, kWave));
    ly = box_min(ly, __shfl_xor(ly, o, kWave));
    hx = box_max(hx, __shfl_xor(hx, o, kWave));
    hy = box_max(hy, __shfl_xor(hy, o, kWave));
  }
  if (lane_id() == 0) {
    minx[row] = lx;
    miny[row] = ly;
    maxx[row] = hx;
    maxy[row] = hy;
  }
}

}  // namespace

}  // namespace qpip

extern "C" int qpip_polygon_bounding_boxes(const int32_t* poly_offsets,
                                           int64_t num_poly_offsets,
                                           const int32_t* ring_offsets,
                                           int64_t num_ring_offsets,
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
  using namespace qpip;
  try {
    QPIP_EXPECTS(minx && miny && maxx && maxy && num_boxes, "output pointers must not be null");
    *minx = *miny = *maxx = *maxy = nullptr;
    *num_boxes                    = 0;
    hipStream_t stream            = static_cast<hipStream_t>(stream_);
    // polygon_bounding_boxes.cu:140-155 (dtype checks are the binding's: int32 offsets, f64 x/y)
    int64_t const num_polys = num_poly_offsets > 0 ? num_poly_offsets - 1 : 0;
    QPIP_EXPECTS(num_x == num_y, "x and y must be the same size");
    if (num_polys == 0) return QPIP_SUCCESS;
    // detail/bounding_boxes.cuh:165-169 (CUSPATIAL_EXPECTS_VALID_POLYGON_SIZES)
    QPIP_EXPECTS(num_poly_offsets > 0, "Polygon offsets must contain at least one (1) value");
    QPIP_EXPECTS(num_ring_offsets > 0, "Polygon ring offsets must contain at least one (1) value");
    int64_t const num_rings = num_ring_offsets - 1;
    StreamScope scope(stream);
    OutputAllocator oa(mr, stream);
    double* o_minx = oa.make<double>(num_polys);
    double* o_miny = oa.make<double>(num_polys);
    double* o_maxx = oa.make<double>(num_polys);
    double* o_maxy = oa.make<double>(num_polys);
    dbuf<int64_t> lo(num_polys), hi(num_polys);
    dbuf<uint32_t> partials(kScanBlocks + 1), flags(2);
    QPIP_HIP(hipMemsetAsync(flags.get(), 0, 2 * sizeof(uint32_t), stream));
    int64_t const nv = (num_rings == 0) ? 0 : num_x;  // detail/bounding_boxes.cuh:171
    if (nv > 0) {
      IntervalOp op{poly_offsets, ring_offsets, num_poly_offsets, num_ring_offsets, nv, num_polys,
                    lo.get(), hi.get(), flags.get() + 1};
      uint64_t const n_int = static_cast<uint64_t>(num_polys) + 2;
      count_scan_emit<IntervalOp, uint32_t, uint32_t>(op, nullptr, n_int, partials.get(),
                                                      flags.get(), stream, scan_blocks_for(n_int),
                                                      "bbox_intervals");
    }
    {
      // one wave per box, in slices of 2^22 boxes (grids within 2^28 work-items)
      KernelTimer kt("bbox_reduce", stream);
      for (int64_t b = 0; b < num_polys; b += int64_t{1} << 22) {
        int64_t const m = std::min<int64_t>(num_polys - b, int64_t{1} << 22);
        bbox_reduce_kernel<<<static_cast<unsigned>((m + 3) / 4), 256, 0, stream>>>(
          x, y, lo.get(), hi.get(), flags.get(), num_polys, expansion_radius, o_minx, o_miny,
          o_maxx, o_maxy, b);
        QPIP_LAUNCH_CHECK();
      }
      QPIP_LAUNCH_CHECK();
    }
    uint32_t h_flags[2];
    QPIP_HIP(hipMemcpyAsync(h_flags, flags.get(), sizeof(h_flags), hipMemcpyDeviceToHost, stream));
    QPIP_HIP(hipStreamSynchronize(stream));
    if (h_flags[1])  // the outputs are released by `oa`
      throw logic_error("cuSpatial failure: polygon offset out of range of ring offsets");
    oa.keep();
    *minx      = o_minx;
    *miny      = o_miny;
    *maxx      = o_maxx;
    *maxy      = o_maxy;
    *num_boxes = num_polys;
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}
