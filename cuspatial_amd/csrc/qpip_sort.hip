// Stable LSD radix sort of (uint32 key, uint32 value) pairs on the low `bits`
// bits of the key. Replaces the thrust::stable_sort_by_key calls of the path
// (phase_1.cuh:92, quadtree_bbox_filtering.cuh:161-165), which sort all 32 bits;
// only the bits that can be non-zero are sorted here.
//
// One pass per 8-bit digit, reduce-then-scan:
//   count   : per-tile digit histogram (skipped for pass 0 when the producer
//             kernel already wrote it, as the fused keygen does);
//   scan    : one workgroup per digit scans that digit's row of tile counts;
//   scatter : tiles of 4096 pairs are ranked in registers with wave64 ballots
//             (stable: wave-striped order is index order), staged through LDS in
//             digit order, and written out as runs of equal digits.
#include <string>

#include "qpip_sort.hpp"

namespace qpip {

namespace {

__device__ __forceinline__ uint32_t digit_of(uint32_t key, int shift, uint32_t mask)
{
  return (key >> shift) & mask;
}

__global__ void __launch_bounds__(kSortThreads)
  radix_count_kernel(uint32_t const* __restrict__ keys, uint64_t n, uint64_t tiles, int shift,
                     uint32_t mask, uint32_t* __restrict__ counts)
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
    uint32_t const d   = valid ? digit_of(keys[idx], shift, mask) : 0u;
    wave_hist_add(hist, d, valid);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) counts[d * tiles + tile] = hist[d];
}

// Block d scans row d of counts (exclusive, in place) and writes its total.
__global__ void __launch_bounds__(256)
  radix_scan_kernel(uint32_t* __restrict__ counts, uint64_t tiles, uint32_t* __restrict__ row_total)
{
  __shared__ uint32_t lds[256 / kWave + 1];
  uint32_t* row = counts + static_cast<uint64_t>(blockIdx.x) * tiles;
  uint32_t base = 0;
  for (uint64_t t0 = 0; t0 < tiles; t0 += 256) {
    uint64_t const i = t0 + threadIdx.x;
    uint32_t const c = i < tiles ? row[i] : 0u;
    uint32_t tot;
    uint32_t const ex = block_excl_scan<256>(c, lds, &tot);
    if (i < tiles) row[i] = base + ex;
    base += tot;
  }
  if (threadIdx.x == 0) row_total[blockIdx.x] = base;
}

template <int NB>
__global__ void __launch_bounds__(kSortThreads)
  radix_scatter_kernel(uint32_t const* __restrict__ keys_in,
                       uint32_t const* __restrict__ vals_in,  // null: identity
                       uint32_t* __restrict__ keys_out,
                       uint32_t* __restrict__ vals_out,
                       uint64_t n,
                       uint64_t tiles,
                       int shift,
                       uint32_t const* __restrict__ counts,     // scanned rows
                       uint32_t const* __restrict__ row_total)  // per-digit totals
{
  constexpr int NW = kSortThreads / kWave;
  __shared__ uint32_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];
  __shared__ uint32_t s_wcount[NW][kRadix];
  __shared__ uint32_t s_tile_start[kRadix];
  __shared__ uint32_t s_global[kRadix];
  __shared__ uint32_t s_scan[kSortThreads / kWave + 1];

  uint32_t const mask = (1u << NB) - 1u;
  uint64_t const tile = blockIdx.x;
  int const w         = threadIdx.x / kWave;
  int const lane      = lane_id();

  for (int i = threadIdx.x; i < NW * kRadix; i += kSortThreads) (&s_wcount[0][0])[i] = 0;
  __syncthreads();

  uint64_t const base = tile * kSortTile + static_cast<uint64_t>(w) * (kSortItems * kWave);
  uint32_t key[kSortItems];
  uint32_t val[kSortItems];
  uint32_t rank[kSortItems];
  uint64_t const lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t const idx = base + static_cast<uint64_t>(j) * kWave + lane;
    bool const valid   = idx < n;
    key[j]             = valid ? keys_in[idx] : 0xFFFFFFFFu;
    val[j]             = valid ? (vals_in ? vals_in[idx] : static_cast<uint32_t>(idx)) : 0u;
    uint32_t const d   = digit_of(key[j], shift, mask);
    uint64_t const act = ballot(valid);
    uint64_t const peers = wave_match_digit<NB>(d, act);
    uint32_t const prior = s_wcount[w][d];
    rank[j]              = prior + static_cast<uint32_t>(__popcll(peers & lt));
    if (valid && lane == __builtin_ctzll(peers))
      s_wcount[w][d] = prior + static_cast<uint32_t>(__popcll(peers));
  }
  __syncthreads();

  // Per digit: exclusive prefix over waves, tile-local digit starts, global base.
  {
    int const d = threadIdx.x;  // kSortThreads == kRadix
    uint32_t acc = 0;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) {
      uint32_t const c = s_wcount[ww][d];
      s_wcount[ww][d]  = acc;
      acc += c;
    }
    uint32_t tot;
    uint32_t const start = block_excl_scan<kSortThreads>(acc, s_scan, &tot);
    s_tile_start[d]      = start;
    uint32_t const rt    = row_total[d];
    uint32_t const dbase = block_excl_scan<kSortThreads>(rt, s_scan, &tot);
    s_global[d]          = dbase + counts[static_cast<uint64_t>(d) * tiles + tile];
  }
  __syncthreads();

#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t const idx = base + static_cast<uint64_t>(j) * kWave + lane;
    if (idx < n) {
      uint32_t const d   = digit_of(key[j], shift, mask);
      uint32_t const loc = s_tile_start[d] + s_wcount[w][d] + rank[j];
      s_keys[loc]        = key[j];
      s_vals[loc]        = val[j];
    }
  }
  __syncthreads();

  uint64_t const tile_lo = tile * kSortTile;
  uint32_t const valid_n =
    static_cast<uint32_t>((n - tile_lo) < static_cast<uint64_t>(kSortTile) ? (n - tile_lo)
                                                                            : kSortTile);
#pragma unroll 4
  for (int k = 0; k < kSortItems; ++k) {
    uint32_t const i = static_cast<uint32_t>(k * kSortThreads + threadIdx.x);
    if (i < valid_n) {
      uint32_t const kk  = s_keys[i];
      uint32_t const d   = digit_of(kk, shift, mask);
      uint64_t const pos = static_cast<uint64_t>(s_global[d]) + (i - s_tile_start[d]);
      keys_out[pos]      = kk;
      vals_out[pos]      = s_vals[i];
    }
  }
}

template <int NB>
void launch_scatter(uint32_t const* ki, uint32_t const* vi, uint32_t* ko, uint32_t* vo, uint64_t n,
                    uint64_t tiles, int shift, uint32_t const* counts, uint32_t const* row_total,
                    hipStream_t stream)
{
  radix_scatter_kernel<NB><<<static_cast<unsigned>(tiles), kSortThreads, 0, stream>>>(
    ki, vi, ko, vo, n, tiles, shift, counts, row_total);
  QPIP_LAUNCH_CHECK();
}

}  // namespace

SortResult radix_sort_pairs(uint32_t* keys_a,
                            uint32_t* keys_b,
                            uint32_t const* values_in,
                            uint32_t* values_a,
                            uint32_t* values_b,
                            uint64_t n,
                            int bits,
                            hipStream_t stream,
                            uint32_t* pass0_counts,
                            char const* tag)
{
  std::string const t_count   = std::string(tag) + "_count";
  std::string const t_scan    = std::string(tag) + "_scan";
  std::string const t_scatter = std::string(tag) + "_scatter";
  static_assert(kSortThreads == kRadix, "one thread per digit in the scatter prologue");
  if (n == 0) return {keys_a, values_in ? const_cast<uint32_t*>(values_in) : values_a};
  if (bits < 1) bits = 1;
  if (bits > 32) bits = 32;
  uint64_t const tiles = sort_tiles(n);
  if (tiles > 0x7FFFFFFFull) throw logic_error("cuSpatial failure: sort input too large");
  // pass0_counts, when given, already holds pass 0's tile histograms (written by
  // the producer of the keys) and serves as the histogram buffer of every pass.
  dbuf<uint32_t> own_counts;
  if (!pass0_counts) own_counts.reset(static_cast<size_t>(kRadix) * tiles);
  uint32_t* const cnt = pass0_counts ? pass0_counts : own_counts.get();
  dbuf<uint32_t> row_total(kRadix);

  uint32_t* ksrc       = keys_a;
  uint32_t* kdst       = keys_b;
  uint32_t const* vsrc = values_in;  // null on the first pass: identity
  uint32_t* vdst       = values_a;
  uint32_t* vother     = values_b;
  int const passes     = (bits + kRadixBits - 1) / kRadixBits;
  for (int p = 0; p < passes; ++p) {
    int const shift = p * kRadixBits;
    int const nb    = (bits - shift) < kRadixBits ? (bits - shift) : kRadixBits;
    if (!(p == 0 && pass0_counts)) {
      KernelTimer kt(t_count.c_str(), stream);
      radix_count_kernel<<<static_cast<unsigned>(tiles), kSortThreads, 0, stream>>>(
        ksrc, n, tiles, shift, (1u << nb) - 1u, cnt);
      QPIP_LAUNCH_CHECK();
    }
    {
      KernelTimer kt(t_scan.c_str(), stream);
      radix_scan_kernel<<<kRadix, 256, 0, stream>>>(cnt, tiles, row_total.get());
      QPIP_LAUNCH_CHECK();
    }
    KernelTimer kt(t_scatter.c_str(), stream);
    switch (nb) {
      case 1: launch_scatter<1>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
      case 2: launch_scatter<2>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
      case 3: launch_scatter<3>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
      case 4: launch_scatter<4>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
      case 5: launch_scatter<5>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
      case 6: launch_scatter<6>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
      case 7: launch_scatter<7>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
      default: launch_scatter<8>(ksrc, vsrc, kdst, vdst, n, tiles, shift, cnt, row_total.get(), stream); break;
    }
    // next pass reads what this one wrote
    std::swap(ksrc, kdst);
    vsrc = vdst;
    std::swap(vdst, vother);
  }
  return {ksrc, const_cast<uint32_t*>(vsrc)};
}

}  // namespace qpip
