// Functor-driven device-wide exclusive scans ("count, scan, emit").
//
// Many stages of the path have the same shape: every element i of a sequence
// contributes count(i) outputs, and element i writes its outputs at the exclusive
// prefix sum of the counts (the reference does this with thrust::exclusive_scan
// followed by scatter / copy_if, e.g. phase_2.cuh:197-208, traversal.cuh:216-240,
// quadtree_point_in_polygon.cuh:145-207). Here it is two launches on a fixed
// grid, whose length may live on the device (so that no host synchronisation is
// needed between data-dependent stages):
//   1. chunk_count: block b reduces the counts of its contiguous chunk;
//   2. chunk_emit:  block b sums the chunk sums of the blocks before it (at most
//                   1024 values, a few KB of L2), rescans its chunk and calls
//                   emit(i, prefix); block 0 also stores the grand total.
// The functor's count() is evaluated twice (passes 1 and 2), which is cheaper
// than storing it when it is a few loads.
#pragma once

#include "qpip_internal.hpp"

namespace qpip {

constexpr int kScanThreads = 256;
constexpr int kScanBlocks  = 1024;  // fixed grid: 4 workgroups per CU on 256 CUs

// Chunk of block b for a sequence of n elements over G blocks; chunks are
// multiples of the block size so that thread t handles elements t, t+256, ...
__device__ __forceinline__ void chunk_range(uint64_t n, uint64_t b, uint64_t G, uint64_t& lo,
                                            uint64_t& hi)
{
  uint64_t per = (n + G - 1) / G;
  per          = (per + kScanThreads - 1) / kScanThreads * kScanThreads;
  lo           = b * per;
  hi           = lo + per;
  if (lo > n) lo = n;
  if (hi > n) hi = n;
}

__device__ __forceinline__ uint64_t load_len(uint64_t const* n_dev, uint64_t n_host)
{
  return n_dev ? *n_dev : n_host;
}

template <class Op, class C>
__global__ void __launch_bounds__(kScanThreads) chunk_count_kernel(Op op,
                                                                   uint64_t const* n_dev,
                                                                   uint64_t n_host,
                                                                   C* partials)
{
  __shared__ C lds[kScanThreads / kWave + 1];
  uint64_t const n = load_len(n_dev, n_host);
  uint64_t lo, hi;
  chunk_range(n, blockIdx.x, gridDim.x, lo, hi);
  C acc = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += kScanThreads) acc += op.count(i);
  acc = wave_sum(acc);
  if (lane_id() == 0) lds[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    C s = 0;
    for (int w = 0; w < kScanThreads / kWave; ++w) s += lds[w];
    partials[blockIdx.x] = s;
  }
}

template <class Op, class C, class TT>
__global__ void __launch_bounds__(kScanThreads) chunk_emit_kernel(Op op,
                                                                  uint64_t const* n_dev,
                                                                  uint64_t n_host,
                                                                  C* partials,
                                                                  TT* total_dev)
{
  __shared__ C lds[kScanThreads / kWave + 1];
  uint64_t const n = load_len(n_dev, n_host);
  uint64_t lo, hi;
  chunk_range(n, blockIdx.x, gridDim.x, lo, hi);
  C before = 0, all = 0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += kScanThreads) {
    C const p = partials[i];
    all += p;
    if (i < blockIdx.x) before += p;
  }
  C base, total;
  block_excl_scan<kScanThreads>(before, lds, &base);
  block_excl_scan<kScanThreads>(all, lds, &total);
  if (blockIdx.x == 0 && threadIdx.x == 0 && total_dev) *total_dev = total;  // TT converts
  for (uint64_t t0 = lo; t0 < hi; t0 += kScanThreads) {
    uint64_t const i = t0 + threadIdx.x;
    C const c        = i < hi ? op.count(i) : C(0);
    C tile_total;
    C const ex = block_excl_scan<kScanThreads>(c, lds, &tile_total);
    if (i < hi) op.emit(i, base + ex);
    base += tile_total;
  }
}

// Runs count/scan/emit over a sequence whose length is *n_dev (device) or
// n_host (when n_dev is null). `partials` needs `blocks` entries; the total is,
// if given, converted into *total_dev.
template <class Op, class C, class TT = C>
void count_scan_emit(Op const& op,
                     uint64_t const* n_dev,
                     uint64_t n_host,
                     C* partials,
                     TT* total_dev,
                     hipStream_t stream,
                     int blocks          = kScanBlocks,
                     char const* timer   = "scan")
{
  if (blocks > kScanBlocks) blocks = kScanBlocks;
  KernelTimer kt(timer, stream);
  chunk_count_kernel<Op, C><<<blocks, kScanThreads, 0, stream>>>(op, n_dev, n_host, partials);
  QPIP_LAUNCH_CHECK();
  chunk_emit_kernel<Op, C, TT>
    <<<blocks, kScanThreads, 0, stream>>>(op, n_dev, n_host, partials, total_dev);
  QPIP_LAUNCH_CHECK();
}

// Blocks to use for a host-known length: enough to cover n with one element
// per thread, at most kScanBlocks.
inline int scan_blocks_for(uint64_t n)
{
  uint64_t b = (n + kScanThreads - 1) / kScanThreads;
  if (b < 1) b = 1;
  if (b > static_cast<uint64_t>(kScanBlocks)) b = kScanBlocks;
  return static_cast<int>(b);
}

}  // namespace qpip
