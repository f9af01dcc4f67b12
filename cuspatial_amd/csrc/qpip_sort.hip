// Stable LSD radix sort of (uint32 key, uint32 value) pairs on the low `bits`
// bits of the key. Replaces the thrust::stable_sort_by_key calls of the path
// (phase_1.cuh:92, quadtree_bbox_filtering.cuh:161-165), which sort all 32 bits;
// only the bits that can be non-zero are sorted here.
//
// One 8-bit digit per pass, reduce-then-scan over tiles of 4096 pairs. No block
// ever waits for another, so nothing depends on the order in which workgroups are
// dispatched:
//   count:    one block per tile counts its digits (LDS atomics into per-wave
//             histograms) -> counts[tile][256];
//   colscan:  counts[tile][d] becomes the first output position of the tile's
//             digit-d pairs: digit base (exclusive scan of the digit totals) + the
//             digit's count in earlier tiles. Three small kernels over groups of
//             kScanRows tiles (reduce, scan of the group sums, rescan);
//   scatter:  one block per tile (4096 pairs, 16 per thread, wave-striped so that
//             register order is index order) ranks the tile with wave64 ballot peer
//             matching, stages it through LDS in digit order, and writes each
//             digit's run at its position.
// The first pass's counts usually come from the producer of the keys (the keygen
// kernel counts digit 0 while it writes the keys).
#include <algorithm>
#include <string>
#include <vector>

#include "qpip_sort.hpp"

namespace qpip {

namespace {

__device__ __forceinline__ uint32_t digit_of(uint32_t key, int shift, uint32_t mask)
{
  return (key >> shift) & mask;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

constexpr int kScanRows = 64;  // tiles per column-scan group
constexpr int kBatch    = 8;   // loads in flight per thread in the column scans

// Digit counts of one tile. `vec`: the keys are 16-byte aligned (full tiles are
// then read as uint4).
__global__ void __launch_bounds__(kSortThreads)
  radix_count_kernel(uint32_t const* __restrict__ keys, uint64_t n, int shift, uint32_t mask,
                     bool vec, uint32_t* __restrict__ counts)
{
  constexpr int NW = kSortThreads / kWave;
  __shared__ uint32_t h[NW][kRadix];
  int const w = threadIdx.x / kWave;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) h[ww][threadIdx.x] = 0;
  __syncthreads();
  uint64_t const lo = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  if (vec && lo + kSortTile <= n) {
    uint4 const* k4 = reinterpret_cast<uint4 const*>(keys + lo);
    uint4 v[kSortItems / 4];
#pragma unroll
    for (int j = 0; j < kSortItems / 4; ++j) v[j] = k4[j * kSortThreads + threadIdx.x];
#pragma unroll
    for (int j = 0; j < kSortItems / 4; ++j) {
      wave_count_digit(h[w], digit_of(v[j].x, shift, mask));
      wave_count_digit(h[w], digit_of(v[j].y, shift, mask));
      wave_count_digit(h[w], digit_of(v[j].z, shift, mask));
      wave_count_digit(h[w], digit_of(v[j].w, shift, mask));
    }
  } else {
    uint32_t d[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
      uint64_t const idx = lo + static_cast<uint64_t>(j) * kSortThreads + threadIdx.x;
      d[j]               = idx < n ? digit_of(keys[idx], shift, mask) : kRadix;
    }
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) wave_count_digit(h[w], d[j]);
  }
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) s += h[ww][threadIdx.x];
  counts[static_cast<uint64_t>(blockIdx.x) * kRadix + threadIdx.x] = s;
}

// Column sums of counts over the tiles of group g (rows [g * kScanRows, ...)).
__global__ void __launch_bounds__(kRadix)
  colscan_reduce_kernel(uint32_t const* __restrict__ counts, uint64_t tiles,
                        uint32_t* __restrict__ sums)
{
  int const d       = threadIdx.x;
  uint64_t const r0 = static_cast<uint64_t>(blockIdx.x) * kScanRows;
  uint64_t const r1 = r0 + kScanRows < tiles ? r0 + kScanRows : tiles;
  uint32_t s        = 0;
  for (uint64_t r = r0; r < r1; r += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = r + u < r1 ? counts[(r + u) * kRadix + d] : 0u;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) s += v[u];
  }
  sums[static_cast<uint64_t>(blockIdx.x) * kRadix + d] = s;
}

// sums[g][d] (in place) -> first output position of digit d's pairs of group g:
// digit base (exclusive scan of the digit totals, in digit order) + the digit's
// count in earlier groups. 1024 threads: digit d = threadIdx % 256 over quarter
// q = threadIdx / 256 of the groups.
__global__ void __launch_bounds__(1024)
  colscan_top_kernel(uint32_t* __restrict__ sums, uint64_t groups)
{
  __shared__ uint32_t qsum[4][kRadix];
  __shared__ uint32_t dbase[kRadix];
  __shared__ uint32_t lds[1024 / kWave + 1];
  int const d        = threadIdx.x % kRadix;
  int const q        = threadIdx.x / kRadix;
  uint64_t const per = (groups + 3) / 4;
  uint64_t g0        = static_cast<uint64_t>(q) * per;
  uint64_t g1        = g0 + per;
  if (g0 > groups) g0 = groups;
  if (g1 > groups) g1 = groups;
  uint32_t s = 0;
  for (uint64_t g = g0; g < g1; g += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = g + u < g1 ? sums[(g + u) * kRadix + d] : 0u;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) s += v[u];
  }
  qsum[q][d] = s;
  __syncthreads();
  uint32_t total = 0;
  if (q == 0) total = qsum[0][d] + qsum[1][d] + qsum[2][d] + qsum[3][d];
  uint32_t tot;
  uint32_t const base = block_excl_scan<1024>(total, lds, &tot);
  if (q == 0) dbase[d] = base;
  __syncthreads();
  uint32_t run = dbase[d];
  for (int k = 0; k < q; ++k) run += qsum[k][d];
  for (uint64_t g = g0; g < g1; g += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = g + u < g1 ? sums[(g + u) * kRadix + d] : 0u;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (g + u < g1) sums[(g + u) * kRadix + d] = run;
      run += v[u];
    }
  }
}

// counts[t][d] (in place) -> first output position of tile t's digit-d pairs.
__global__ void __launch_bounds__(kRadix)
  colscan_down_kernel(uint32_t* __restrict__ counts, uint64_t tiles,
                      uint32_t const* __restrict__ sums)
{
  int const d       = threadIdx.x;
  uint64_t const r0 = static_cast<uint64_t>(blockIdx.x) * kScanRows;
  uint64_t const r1 = r0 + kScanRows < tiles ? r0 + kScanRows : tiles;
  uint32_t run      = sums[static_cast<uint64_t>(blockIdx.x) * kRadix + d];
  for (uint64_t r = r0; r < r1; r += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = r + u < r1 ? counts[(r + u) * kRadix + d] : 0u;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (r + u < r1) counts[(r + u) * kRadix + d] = run;
      run += v[u];
    }
  }
}

// Ranks one tile: keys/values in registers, per-wave digit counts in s_wcount_w,
// rank[j] = position of item j among equal digits within its wave (stable).
template <int NB>
__device__ __forceinline__ void rank_tile(uint32_t const* __restrict__ keys_in,
                                          uint32_t const* __restrict__ vals_in, uint64_t n,
                                          uint64_t base, int shift, uint32_t (&key)[kSortItems],
                                          uint32_t (&val)[kSortItems],
                                          uint32_t (&rank)[kSortItems], uint32_t* s_wcount_w)
{
  uint32_t const mask = (1u << NB) - 1u;
  int const lane      = lane_id();
  uint64_t const lt   = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t const idx = base + static_cast<uint64_t>(j) * kWave + lane;
    bool const valid   = idx < n;
    key[j]             = valid ? keys_in[idx] : 0xFFFFFFFFu;
    val[j]             = valid ? (vals_in ? vals_in[idx] : static_cast<uint32_t>(idx)) : 0u;
  }
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t const idx   = base + static_cast<uint64_t>(j) * kWave + lane;
    bool const valid     = idx < n;
    uint32_t const d     = digit_of(key[j], shift, mask);
    uint64_t const act   = ballot(valid);
    uint64_t const peers = wave_match_digit<NB>(d, act);
    uint32_t const prior = s_wcount_w[d];
    rank[j]              = prior + static_cast<uint32_t>(__popcll(peers & lt));
    if (valid && lane == __builtin_ctzll(peers))
      s_wcount_w[d] = prior + static_cast<uint32_t>(__popcll(peers));
  }
}

// Workgroups are dealt to the 8 XCDs round-robin (block b to XCD b % 8). Giving
// XCD x the x-th eighth of the tiles, in order, makes the tiles that write the
// neighbouring pieces of each digit's runs share an L2, where partial lines merge.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t b, uint64_t tiles)
{
  uint64_t const per = tiles / 8, rem = tiles % 8, x = b % 8, r = b / 8;
  return x * per + (x < rem ? x : rem) + r;
}

// SEG: the tile is entry `tile` of a segment tile table (its own start and count,
// never straddling two segments), and with `compose` the keys are written OR-ed with
// the segment's key_or and, when `vals_local` is given, each value minus its
// segment's first position is written there too (with `local_out`, in place of
// the value itself); `key_samples`, when given, receives the key written at
// every position that is a multiple of 32.
template <int NB, bool SEG>
__global__ void __launch_bounds__(kSortThreads)
  radix_scatter_kernel(uint32_t const* __restrict__ keys_in,
                       uint32_t const* __restrict__ vals_in,  // null: identity
                       uint32_t* __restrict__ keys_out,
                       uint32_t* __restrict__ vals_out,
                       uint64_t n,
                       int shift,
                       uint32_t const* __restrict__ offsets,  // tiles x 256, from the colscan
                       SortTile const* __restrict__ seg_tiles,
                       SortSegment const* __restrict__ segments,
                       bool compose,
                       uint32_t* __restrict__ vals_local,
                       uint32_t* __restrict__ key_samples,
                       bool local_out)
{
  constexpr int NW = kSortThreads / kWave;
  __shared__ uint32_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];
  __shared__ uint32_t s_wcount[NW][kRadix];
  __shared__ uint32_t s_tile_start[kRadix];
  __shared__ uint32_t s_global[kRadix];
  __shared__ uint32_t s_scan[kSortThreads / kWave + 1];

  uint32_t const mask = (1u << NB) - 1u;
  int const w         = threadIdx.x / kWave;
  int const lane      = lane_id();
  int const d         = threadIdx.x;  // kSortThreads == kRadix: one digit per thread
  uint64_t const tile = xcd_tile(blockIdx.x, gridDim.x);
  s_global[d]         = offsets[tile * kRadix + d];
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) s_wcount[ww][d] = 0;
  __syncthreads();
  uint64_t tile_lo = tile * kSortTile, end = n;
  uint32_t key_or = 0, seg_begin = 0;
  if constexpr (SEG) {
    SortTile const T = seg_tiles[tile];
    tile_lo          = T.start;
    end              = static_cast<uint64_t>(T.start) + T.count;
    if (compose) {
      key_or    = segments[T.segment].key_or;
      seg_begin = segments[T.segment].begin;
    }
  }
  uint64_t const base = tile_lo + static_cast<uint64_t>(w) * (kSortItems * kWave);
  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
  rank_tile<NB>(keys_in, vals_in, end, base, shift, key, val, rank, s_wcount[w]);
  __syncthreads();
  // per digit: exclusive prefix over waves and the tile's first slot of the digit
  uint32_t count = 0;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) {
    uint32_t const c = s_wcount[ww][d];
    s_wcount[ww][d]  = count;
    count += c;
  }
  uint32_t tot;
  s_tile_start[d] = block_excl_scan<kSortThreads>(count, s_scan, &tot);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t const idx = base + static_cast<uint64_t>(j) * kWave + lane;
    if (idx < end) {
      uint32_t const dj  = digit_of(key[j], shift, mask);
      uint32_t const loc = s_tile_start[dj] + s_wcount[w][dj] + rank[j];
      s_keys[loc]        = key[j];
      s_vals[loc]        = val[j];
    }
  }
  __syncthreads();
  uint32_t const valid_n =
    static_cast<uint32_t>((end - tile_lo) < static_cast<uint64_t>(kSortTile) ? (end - tile_lo)
                                                                              : kSortTile);
#pragma unroll 4
  for (int k = 0; k < kSortItems; ++k) {
    uint32_t const i = static_cast<uint32_t>(k * kSortThreads + threadIdx.x);
    if (i < valid_n) {
      uint32_t const kk  = s_keys[i];
      uint32_t const dk  = digit_of(kk, shift, mask);
      uint64_t const pos = static_cast<uint64_t>(s_global[dk]) + (i - s_tile_start[dk]);
      uint32_t const v  = s_vals[i];
      uint32_t const ko = SEG ? (kk | key_or) : kk;
      keys_out[pos]      = ko;
      vals_out[pos]      = (SEG && local_out) ? v - seg_begin : v;
      if (SEG && vals_local) vals_local[pos] = v - seg_begin;
      if (SEG && key_samples && (pos & 31u) == 0) key_samples[pos >> 5] = ko;
    }
  }
}

// Digit counts of tile `blockIdx.x` of a segment tile table.
__global__ void __launch_bounds__(kSortThreads)
  seg_count_kernel(uint32_t const* __restrict__ keys, SortTile const* __restrict__ seg_tiles,
                   int shift, uint32_t mask, uint32_t* __restrict__ counts)
{
  constexpr int NW = kSortThreads / kWave;
  __shared__ uint32_t h[NW][kRadix];
  int const w = threadIdx.x / kWave;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) h[ww][threadIdx.x] = 0;
  __syncthreads();
  SortTile const T = seg_tiles[blockIdx.x];
  uint32_t d[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    uint32_t const i = static_cast<uint32_t>(j) * kSortThreads + threadIdx.x;
    d[j]             = i < T.count ? digit_of(keys[T.start + i], shift, mask) : kRadix;
  }
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) wave_count_digit(h[w], d[j]);
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) s += h[ww][threadIdx.x];
  counts[static_cast<uint64_t>(blockIdx.x) * kRadix + threadIdx.x] = s;
}

// One block per segment: counts[t][d] (in place) -> first output position of tile
// t's digit-d pairs: the segment's begin + the digit base within the segment
// (exclusive scan of its digit totals) + the digit's count in the segment's
// earlier tiles.
__global__ void __launch_bounds__(kRadix)
  seg_colscan_kernel(uint32_t* __restrict__ counts, SortSegment const* __restrict__ segments)
{
  __shared__ uint32_t lds[kRadix / kWave + 1];
  SortSegment const S = segments[blockIdx.x];
  int const d         = threadIdx.x;
  uint32_t total      = 0;
  for (uint32_t t = S.tile_begin; t < S.tile_end; t += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      v[u] = t + u < S.tile_end ? counts[static_cast<uint64_t>(t + u) * kRadix + d] : 0u;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) total += v[u];
  }
  uint32_t tot;
  uint32_t run = S.begin + block_excl_scan<kRadix>(total, lds, &tot);
  for (uint32_t t = S.tile_begin; t < S.tile_end; t += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      v[u] = t + u < S.tile_end ? counts[static_cast<uint64_t>(t + u) * kRadix + d] : 0u;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (t + u < S.tile_end) counts[static_cast<uint64_t>(t + u) * kRadix + d] = run;
      run += v[u];
    }
  }
}

#define QPIP_DISPATCH_NB(nb, CALL) \
  switch (nb) {                    \
    case 1: CALL(1); break;        \
    case 2: CALL(2); break;        \
    case 3: CALL(3); break;        \
    case 4: CALL(4); break;        \
    case 5: CALL(5); break;        \
    case 6: CALL(6); break;        \
    case 7: CALL(7); break;        \
    default: CALL(8); break;       \
  }

}  // namespace

int radix_passes(int bits)
{
  if (bits < 1) bits = 1;
  if (bits > 32) bits = 32;
  return (bits + kRadixBits - 1) / kRadixBits;
}

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream)
{
  if (n == 0) return;
  fill_u32_kernel<<<grid_for(n, 256, 2048), 256, 0, stream>>>(p, n, v);
  QPIP_LAUNCH_CHECK();
}

SortResult radix_sort_pairs(uint32_t* keys_a,
                            uint32_t* keys_b,
                            uint32_t const* values_in,
                            uint32_t* values_out,
                            uint32_t* values_tmp,
                            uint64_t n,
                            int bits,
                            hipStream_t stream,
                            uint32_t* first_counts,
                            char const* tag,
                            uint32_t* counts_ws,
                            uint32_t* sums_ws)
{
  static_assert(kSortThreads == kRadix, "one thread per digit in the scatter prologue");
  if (bits < 1) bits = 1;
  if (bits > 32) bits = 32;
  int const passes = radix_passes(bits);
  if (n == 0) return {keys_a, values_out};
  if (n >= (uint64_t{1} << 32)) throw logic_error("cuSpatial failure: sort input too large");
  uint64_t const tiles         = sort_tiles(n);
  uint64_t const groups        = (tiles + kScanRows - 1) / kScanRows;
  if (tiles > 0x7FFFFFFFull) throw logic_error("cuSpatial failure: sort input too large");
  std::string const t_count   = std::string(tag) + "_count";
  std::string const t_scan    = std::string(tag) + "_colscan";
  std::string const t_scatter = std::string(tag) + "_scatter";

  // the last pass writes values_out
  auto vdst_of = [&](int p) { return ((passes - 1 - p) % 2 == 0) ? values_out : values_tmp; };
  uint32_t* ksrc       = keys_a;
  uint32_t* kdst       = keys_b;
  uint32_t const* vsrc = values_in;  // null on the first pass: identity
  bool const vec = (reinterpret_cast<uintptr_t>(keys_a) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(keys_b) % 16 == 0);

  static_assert(kScanRows == 64, "sort_sums_elems assumes 64 tiles per column-scan group");
  dbuf<uint32_t> own_counts(first_counts || counts_ws ? 0 : static_cast<size_t>(tiles) * kRadix);
  uint32_t* counts = first_counts ? first_counts : counts_ws ? counts_ws : own_counts.get();
  dbuf<uint32_t> own_sums(sums_ws ? 0 : static_cast<size_t>(groups) * kRadix);
  uint32_t* const sums = sums_ws ? sums_ws : own_sums.get();
  for (int p = 0; p < passes; ++p) {
    int const shift     = p * kRadixBits;
    int const nb        = (bits - shift) < kRadixBits ? (bits - shift) : kRadixBits;
    uint32_t const mask = (1u << nb) - 1u;
    uint32_t* vdst      = vdst_of(p);
    if (!(p == 0 && first_counts)) {
      KernelTimer kt(t_count.c_str(), stream);
      radix_count_kernel<<<static_cast<unsigned>(tiles), kSortThreads, 0, stream>>>(
        ksrc, n, shift, mask, vec, counts);
      QPIP_LAUNCH_CHECK();
    }
    {
      KernelTimer kt(t_scan.c_str(), stream);
      colscan_reduce_kernel<<<static_cast<unsigned>(groups), kRadix, 0, stream>>>(counts, tiles,
                                                                               sums);
      QPIP_LAUNCH_CHECK();
      colscan_top_kernel<<<1, 1024, 0, stream>>>(sums, groups);
      QPIP_LAUNCH_CHECK();
      colscan_down_kernel<<<static_cast<unsigned>(groups), kRadix, 0, stream>>>(counts, tiles,
                                                                             sums);
      QPIP_LAUNCH_CHECK();
    }
    {
      KernelTimer kt(t_scatter.c_str(), stream);
#define QPIP_SCATTER(NB)                                                                      \
  radix_scatter_kernel<NB, false><<<static_cast<unsigned>(tiles), kSortThreads, 0, stream>>>( \
    ksrc, vsrc, kdst, vdst, n, shift, counts, nullptr, nullptr, false, nullptr, nullptr, false)
      QPIP_DISPATCH_NB(nb, QPIP_SCATTER);
#undef QPIP_SCATTER
      QPIP_LAUNCH_CHECK();
    }
    std::swap(ksrc, kdst);
    vsrc = vdst;
  }
  return {ksrc, values_out};
}

void make_sort_tiles(std::vector<uint32_t> const& begin, std::vector<SortTile>& tiles,
                     std::vector<SortSegment>& segments, std::vector<uint32_t> const& key_or)
{
  size_t const S = begin.size() - 1;
  tiles.clear();
  segments.resize(S);
  for (size_t s = 0; s < S; ++s) {
    SortSegment& g = segments[s];
    g.begin        = begin[s];
    g.tile_begin   = static_cast<uint32_t>(tiles.size());
    g.key_or       = key_or[s];
    for (uint64_t p = begin[s]; p < begin[s + 1]; p += kSortTile) {
      uint64_t const c = std::min<uint64_t>(kSortTile, begin[s + 1] - p);
      tiles.push_back(SortTile{static_cast<uint32_t>(p), static_cast<uint32_t>(c),
                               static_cast<uint32_t>(s), 0u});
    }
    g.tile_end = static_cast<uint32_t>(tiles.size());
  }
}

SortResult segmented_radix_sort_pairs(uint32_t* keys_a,
                                      uint32_t* keys_b,
                                      uint32_t* values_out,
                                      uint32_t* values_tmp,
                                      uint64_t n,
                                      int bits,
                                      SortTile const* tiles,
                                      uint64_t num_tiles,
                                      SortSegment const* segments,
                                      uint64_t num_segments,
                                      uint32_t* first_counts,
                                      uint32_t* values_local,
                                      uint32_t* key_samples,
                                      bool values_job_local,
                                      hipStream_t stream,
                                      char const* tag)
{
  if (bits < 1) bits = 1;
  if (bits > 32) bits = 32;
  int const passes = radix_passes(bits);
  if (n == 0 || num_tiles == 0) return {keys_a, values_out};
  if (n >= (uint64_t{1} << 32) || num_tiles > 0x7FFFFFFFull)
    throw logic_error("cuSpatial failure: sort input too large");
  std::string const t_count   = std::string(tag) + "_count";
  std::string const t_scan    = std::string(tag) + "_colscan";
  std::string const t_scatter = std::string(tag) + "_scatter";
  auto vdst_of = [&](int p) { return ((passes - 1 - p) % 2 == 0) ? values_out : values_tmp; };
  uint32_t* ksrc       = keys_a;
  uint32_t* kdst       = keys_b;
  uint32_t const* vsrc = nullptr;  // the first pass's values are the positions
  for (int p = 0; p < passes; ++p) {
    int const shift     = p * kRadixBits;
    int const nb        = (bits - shift) < kRadixBits ? (bits - shift) : kRadixBits;
    uint32_t const mask = (1u << nb) - 1u;
    bool const last     = p + 1 == passes;
    uint32_t* vdst      = vdst_of(p);
    if (p > 0) {
      KernelTimer kt(t_count.c_str(), stream);
      seg_count_kernel<<<static_cast<unsigned>(num_tiles), kSortThreads, 0, stream>>>(
        ksrc, tiles, shift, mask, first_counts);
      QPIP_LAUNCH_CHECK();
    }
    {
      KernelTimer kt(t_scan.c_str(), stream);
      seg_colscan_kernel<<<static_cast<unsigned>(num_segments), kRadix, 0, stream>>>(
        first_counts, segments);
      QPIP_LAUNCH_CHECK();
    }
    {
      KernelTimer kt(t_scatter.c_str(), stream);
#define QPIP_SCATTER(NB)                                                                         \
  radix_scatter_kernel<NB, true><<<static_cast<unsigned>(num_tiles), kSortThreads, 0, stream>>>( \
    ksrc, vsrc, kdst, vdst, n, shift, first_counts, tiles, segments, last,               \
    last ? values_local : nullptr, last ? key_samples : nullptr, last && values_job_local)
      QPIP_DISPATCH_NB(nb, QPIP_SCATTER);
#undef QPIP_SCATTER
      QPIP_LAUNCH_CHECK();
    }
    std::swap(ksrc, kdst);
    vsrc = vdst;
  }
  return {ksrc, values_out};
}

}  // namespace qpip
