// Batched quadtree point-in-polygon: many independent jobs (points, polygons,
// bounding box, scale, depth, max_size each), every stage run once for all of them.
//
// Per job the result is exactly what the single-job sequence returns:
// quadtree_on_points -> polygon_bounding_boxes -> join_quadtree_and_bounding_boxes ->
// quadtree_point_in_polygon (with the Python layer's bbox ordering and scale clamp,
// python/cuspatial/cuspatial/core/spatial/indexing.py:165-189, join.py:144-156):
// the job's point_indices, and its (polygon_index, point_index) hits in the same
// order, with job-local indices.
//
// The jobs are processed as one forest in a shared index space (a group of jobs
// whose job number and Morton key fit one 32-bit sort key):
//   * each job's points are sorted on their own Morton keys (a segmented radix
//     sort: tiles never straddle two jobs, digit offsets restart at every job), so
//     every job's points stay in its own input range, in its own key order; the
//     last pass writes the composite key (job << key_bits) | Morton key that the
//     forest's searches use;
//   * level 0 is found per job; the level expansion is the single-tree kernel
//     with per-node split rules (the job's depth and max_size);
//   * the join is the single-tree breadth-first traversal, its top frontier built
//     per job (node = i % T_j, box = i / T_j within the job) and node geometry
//     taken from the node's job; the leaf pairs are sorted by leaf offset, which
//     orders them by job and, within a job, as the single-job join does;
//   * the point-in-polygon test runs on the forest's arrays as one call;
//   * hits and point_indices are translated back to job-local indices.
// A group's jobs never interact: every search is confined to its job's key range.
#include <algorithm>
#include <memory>
#include <vector>

#include "qpip_box.hpp"
#include "qpip_join.hpp"
#include "qpip_pip.hpp"
#include "qpip_sort.hpp"
#include "qpip_tree.hpp"

namespace qpip {

namespace {

constexpr int kBatchMarked    = 16;
constexpr int kMaxLevelsBatch = 16;

// One job of a group, after the reference's clamps (detail/point_quadtree.cuh:259-268).
struct JobDev {
  double min_x, min_y, max_x, max_y, scale;
  uint32_t point_begin, point_end;  // group-local point (and sorted position) range
  uint32_t poly_begin, poly_end;    // group-local polygon range
  uint32_t outside_key;
  int32_t d;
  uint32_t max_size;
  int32_t shift0;
};

// Per job, filled on the device.
struct JobTree {
  uint32_t num_top;     // level-0 nodes (T)
  uint32_t top_first;   // first level-0 node in the forest
  uint32_t num_marked;  // E = R - 1 (root-level quirk)
  uint32_t marked_key[kBatchMarked];
  uint32_t marked_internal[kBatchMarked];
  uint32_t marked_start[kBatchMarked];
  uint32_t marked_count[kBatchMarked];
};

struct ForestMeta {
  uint64_t sizes[kMaxLevelsBatch + 1];
  uint32_t overflow;
  uint64_t num_nodes;
};

__device__ __forceinline__ uint32_t find_job(JobDev const* jobs, uint32_t lo, uint32_t hi,
                                             uint32_t pos)
{
  // last job in [lo, hi] whose point_begin <= pos
  while (lo < hi) {
    uint32_t const mid = lo + ((hi - lo + 1) >> 1);
    if (jobs[mid].point_begin <= pos)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// find_job by a whole wave (every lane calls it; all get the answer): each round
// probes 64 evenly spaced jobs of [lo, hi] at once and keeps the stretch between
// the last probe whose point_begin <= pos and the next probe, so 4096 jobs take 2
// or 3 dependent rounds instead of 12. point_begin never decreases with the job
// index, so the probes' predicate holds for a prefix of the lanes.
__device__ __forceinline__ uint32_t wave_find_job(JobDev const* jobs, uint32_t lo, uint32_t hi,
                                                  uint32_t pos)
{
  int const lane = lane_id();
  while (lo < hi) {
    uint64_t const span = hi - lo;
    // probe(l) = lo + ceil((l + 1) * span / 64): non-decreasing in l, probe(63) = hi
    auto probe = [&](int l) {
      return lo + static_cast<uint32_t>((static_cast<uint64_t>(l + 1) * span + kWave - 1) / kWave);
    };
    bool const ok = jobs[probe(lane)].point_begin <= pos;
    int const c   = __popcll(ballot(ok));  // lanes [0, c) hold
    if (c == 0) {
      hi = probe(0) - 1;
    } else {
      uint32_t const nlo = probe(c - 1);
      hi                 = c < kWave ? probe(c) - 1 : hi;
      lo                 = nlo;
    }
  }
  return lo;
}

// --- keys ----------------------------------------------------------------------------
// Keys of the segmented point sort: one block per tile of the group's tile table
// (a tile never holds two jobs' points), the Morton key alone (the sort's last pass
// adds the job number above it). Also counts the sort's first digit.
__global__ void __launch_bounds__(256)
  batch_keygen_seg_kernel(double const* __restrict__ xy, SortTile const* __restrict__ tiles,
                          JobDev const* __restrict__ jobs, uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ counts)
{
  constexpr int NW = 256 / kWave;
  __shared__ uint32_t h[NW][kRadix];
  int const w = threadIdx.x / kWave;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) h[ww][threadIdx.x] = 0;
  __syncthreads();
  SortTile const T = tiles[blockIdx.x];
  JobDev const J   = jobs[T.segment];
#pragma unroll 4
  for (int k = 0; k < kSortItems; ++k) {
    uint32_t const i = static_cast<uint32_t>(k) * 256 + threadIdx.x;
    uint32_t digit   = kRadix;
    if (i < T.count) {
      uint64_t const p = static_cast<uint64_t>(T.start) + i;
      uint32_t const m = point_key(xy[2 * p], xy[2 * p + 1], J.min_x, J.min_y, J.max_x, J.max_y,
                                   J.scale, J.outside_key);
      keys[p] = m;
      digit   = m & 0xFFu;
    }
    wave_count_digit(h[w], digit);
  }
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) s += h[ww][threadIdx.x];
  counts[static_cast<uint64_t>(blockIdx.x) * kRadix + threadIdx.x] = s;
}

// --- level 0 and the root-level quirk, one wave per job -----------------------------
__global__ void __launch_bounds__(256)
  batch_level0_kernel(uint32_t const* __restrict__ K, JobDev const* __restrict__ jobs,
                      uint32_t num_jobs, int key_bits, uint4* __restrict__ staging,
                      JobTree* __restrict__ trees)
{
  uint32_t const j = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (j >= num_jobs) return;
  JobDev const J     = jobs[j];
  int const v        = lane_id();
  uint32_t const pb  = J.point_begin, pe = J.point_end;
  uint64_t const pre = static_cast<uint64_t>(j) << key_bits;
  // Prefix v covers keys [v << shift0, (v + 1) << shift0) of the job; prefixes past
  // the job's key range (which would reach into the job number bits) hold nothing.
  uint64_t const span = uint64_t{1} << key_bits;
  uint64_t const v0   = static_cast<uint64_t>(v) << J.shift0;
  bool present        = false;
  uint32_t lo = pe, hi = pe;
  if (v0 < span) {
    uint64_t const v1 = static_cast<uint64_t>(v + 1) << J.shift0;
    lo      = lower_bound_u32(K, pb, pe, pre + v0);
    hi      = lower_bound_u32(K, lo, pe, pre + (v1 < span ? v1 : span));
    present = hi > lo;
  }
  uint64_t const m = ballot(present);
  if (present) {
    uint32_t const slot = static_cast<uint32_t>(__popcll(m & lanemask_lt()));
    staging[static_cast<uint64_t>(j) * 64 + slot] = make_uint4(static_cast<uint32_t>(v), lo, hi - lo, 0);
  }
  bool root_present = false;
  if (J.d >= 1 && v < 16) {
    uint64_t const r0 = static_cast<uint64_t>(v) << (2 * J.d);
    if (r0 < span) {
      uint64_t const r1  = static_cast<uint64_t>(v + 1) << (2 * J.d);
      uint32_t const rlo = lower_bound_u32(K, pb, pe, pre + r0);
      uint32_t const rhi = lower_bound_u32(K, rlo, pe, pre + (r1 < span ? r1 : span));
      root_present       = rhi > rlo;
    }
  }
  uint64_t const rm = ballot(root_present);
  if (v == 0) {
    JobTree& T       = trees[j];
    T.num_top        = static_cast<uint32_t>(__popcll(m));
    int const R      = __popcll(rm);
    uint32_t const E = (J.d >= 1 && R > 1) ? static_cast<uint32_t>(R - 1) : 0u;
    T.num_marked     = E;
    uint32_t pos     = pb;
    for (uint32_t e = 0; e < E && pos < pe; ++e) {
      uint32_t const k     = K[pos];
      T.marked_key[e]      = k & ((key_bits < 32) ? ((1u << key_bits) - 1u) : 0xFFFFFFFFu);
      T.marked_internal[e] = 0;
      T.marked_start[e]    = 0;
      T.marked_count[e]    = 0;
      pos                  = lower_bound_u32(K, pos, pe, static_cast<uint64_t>(k) + 1);
    }
  }
}

// Copies each job's level-0 nodes into the forest, job after job.
struct Level0Op {
  JobTree* trees;
  uint4 const* staging;
  uint32_t* nkey;
  uint32_t* nstart;
  uint32_t* ncount;
  uint8_t* nlevel;
  uint32_t* njob;
  __device__ uint32_t count(uint64_t j) const { return trees[j].num_top; }
  __device__ void emit(uint64_t j, uint32_t prefix) const
  {
    uint32_t const t = trees[j].num_top;
    trees[j].top_first = prefix;
    for (uint32_t i = 0; i < t; ++i) {
      uint4 const s = staging[j * 64 + i];
      nkey[prefix + i]   = s.x;
      nstart[prefix + i] = s.y;
      ncount[prefix + i] = s.z;
      nlevel[prefix + i] = 0;
      njob[prefix + i]   = static_cast<uint32_t>(j);
    }
  }
};

// Split rule of the forest: the node's job gives its depth, max_size and key prefix.
struct JobLevels {
  uint32_t* njob;
  JobDev const* jobs;
  int key_bits;
  __device__ __forceinline__ bool rule(uint64_t node, int level, uint32_t& ms, int& sh,
                                       uint64_t& prefix) const
  {
    uint32_t const j = njob[node];
    int const d      = jobs[j].d;
    ms               = jobs[j].max_size;
    sh               = 2 * (d - 2 - level);
    prefix           = static_cast<uint64_t>(j) << key_bits;
    return level <= d - 2;
  }
  __device__ __forceinline__ void inherit(uint64_t parent, uint64_t child) const
  {
    njob[child] = njob[parent];
  }
};

// First forest node in [lo, hi) whose (job, key) is not below (j, k).
__device__ __forceinline__ uint64_t lower_bound_job_key(uint32_t const* njob, uint32_t const* nkey,
                                                        uint64_t lo, uint64_t hi, uint32_t j,
                                                        uint32_t k)
{
  while (lo < hi) {
    uint64_t const mid = lo + ((hi - lo) >> 1);
    bool const below   = njob[mid] < j || (njob[mid] == j && nkey[mid] < k);
    if (below)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Per job: the kept extra "parents" of the root-level quirk (bottom level d - 1).
__global__ void __launch_bounds__(256)
  batch_marked_kernel(uint32_t const* __restrict__ nkey, uint32_t const* __restrict__ nstart,
                      uint32_t const* __restrict__ ncount, uint32_t const* __restrict__ njob,
                      JobDev const* __restrict__ jobs, uint32_t num_jobs,
                      ForestMeta const* __restrict__ meta, JobTree* __restrict__ trees)
{
  uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= num_jobs) return;
  JobTree& T = trees[j];
  if (T.num_marked == 0) return;
  int const L = jobs[j].d - 1;  // bottom level (d >= 1 when marked)
  uint64_t b  = 0;
  for (int l = 0; l < L; ++l) b += meta->sizes[l];
  uint64_t const e_ = b + meta->sizes[L];
  for (uint32_t e = 0; e < T.num_marked; ++e) {
    uint64_t const at = lower_bound_job_key(njob, nkey, b, e_, j, T.marked_key[e]);
    if (at < e_ && njob[at] == j && nkey[at] == T.marked_key[e] && ncount[at] > jobs[j].max_size) {
      T.marked_internal[e] = 1;
      T.marked_start[e]    = nstart[at];
      T.marked_count[e]    = ncount[at];
    }
  }
}

// The forest's is_internal / length / offset (detail/point_quadtree.cuh:88-133): an
// internal node's offset is its first child's forest index, a leaf's its first
// point position in the group's sorted order (minus the quirk's skipped points).
__device__ __forceinline__ void finalize_node(
  uint32_t const* __restrict__ nkey, uint32_t const* __restrict__ nstart,
  uint32_t const* __restrict__ ncount, uint8_t const* __restrict__ nlevel,
  uint32_t const* __restrict__ njob, uint32_t const* __restrict__ cfirst,
  uint32_t const* __restrict__ ccount, JobDev const* __restrict__ jobs,
  JobTree const* __restrict__ trees, uint64_t i, uint8_t* __restrict__ internal_out,
  uint32_t* __restrict__ length_out, uint32_t* __restrict__ offset_out,
  uint32_t* __restrict__ base_out)
{
  uint32_t const j     = njob[i];
  JobDev const& J      = jobs[j];
  int const level      = nlevel[i];
  uint32_t const count = ncount[i];
  uint32_t const start = nstart[i];
  bool internal        = level <= J.d - 2 && count > J.max_size;
  uint32_t length      = internal ? ccount[i] : count;
  uint32_t offset      = internal ? cfirst[i] : start;
  JobTree const& T     = trees[j];
  if (T.num_marked > 0 && !internal) {
    bool marked      = false;
    uint32_t skipped = 0;
    for (uint32_t e = 0; e < T.num_marked; ++e) {
      if (!T.marked_internal[e]) continue;
      if (level == J.d - 1 && T.marked_key[e] == nkey[i]) marked = true;
      if (T.marked_start[e] < start) skipped += T.marked_count[e];
    }
    if (marked) {
      internal = true;  // no children: the join never descends into it
      length   = 0;
      offset   = 0;
    } else {
      offset = start - skipped;
    }
  }
  internal_out[i] = internal ? 1 : 0;
  length_out[i]   = length;
  offset_out[i]   = offset;
  if (base_out) base_out[i] = J.point_begin;
}

__global__ void __launch_bounds__(256)
  batch_finalize_kernel(uint32_t const* __restrict__ nkey, uint32_t const* __restrict__ nstart,
                        uint32_t const* __restrict__ ncount, uint8_t const* __restrict__ nlevel,
                        uint32_t const* __restrict__ njob, uint32_t const* __restrict__ cfirst,
                        uint32_t const* __restrict__ ccount, JobDev const* __restrict__ jobs,
                        JobTree const* __restrict__ trees, uint64_t m,
                        ForestMeta const* __restrict__ meta, int levels,
                        uint8_t* __restrict__ internal_out, uint32_t* __restrict__ length_out,
                        uint32_t* __restrict__ offset_out, uint32_t* __restrict__ base_out)
{
  if (meta) {  // node count on the device: the sum of the level sizes
    m = 0;
    for (int l = 0; l < levels; ++l) m += meta->sizes[l];
  }
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    finalize_node(nkey, nstart, ncount, nlevel, njob, cfirst, ccount, jobs, trees, i,
                  internal_out, length_out, offset_out, base_out);
}

// --- join ----------------------------------------------------------------------------
// Node geometry of the forest: the node's job.
struct JobGeometry {
  uint32_t const* njob;
  JobDev const* jobs;
  __device__ __forceinline__ void get(uint32_t node, double& vx, double& vy, double& sc,
                                      int& d) const
  {
    JobDev const& J = jobs[njob[node]];
    vx              = J.min_x;
    vy              = J.min_y;
    sc              = J.scale;
    d               = J.d;
  }
};

// Top frontier sizes T_j * B_j.
struct TopSizeOp {
  JobTree const* trees;
  JobDev const* jobs;
  uint64_t* top_off;
  __device__ uint64_t count(uint64_t j) const
  {
    return static_cast<uint64_t>(trees[j].num_top) * (jobs[j].poly_end - jobs[j].poly_begin);
  }
  __device__ void emit(uint64_t j, uint64_t prefix) const { top_off[j] = prefix; }
};

// Level 0 of the traversal: job j's virtual frontier i -> (i % T_j, i / T_j).
struct BatchTopOp {
  TreeView t;
  BoxView bx;
  JobGeometry g;
  JobTree const* trees;
  JobDev const* jobs;
  uint64_t const* top_off;
  uint32_t num_jobs;
  JoinMeta* meta;
  uint32_t* out_node;
  uint32_t* out_bbox;
  uint32_t* fr_node;
  uint32_t* fr_bbox;
  uint64_t cap_out, cap_fr;

  __device__ void locate(uint64_t i, uint32_t& node, uint32_t& bbox) const
  {
    uint32_t lo = 0, hi = num_jobs - 1;  // last job with top_off <= i
    while (lo < hi) {
      uint32_t const mid = lo + ((hi - lo + 1) >> 1);
      if (top_off[mid] <= i)
        lo = mid;
      else
        hi = mid - 1;
    }
    uint64_t const local = i - top_off[lo];
    uint32_t const T     = trees[lo].num_top;
    node                 = trees[lo].top_first + static_cast<uint32_t>(local % T);
    bbox                 = jobs[lo].poly_begin + static_cast<uint32_t>(local / T);
  }
  __device__ uint64_t count(uint64_t i) const
  {
    uint32_t node, bbox;
    locate(i, node, bbox);
    int const c = classify(t, bx, g, node, bbox);
    return c == 0 ? 1ull : (c == 1 ? (1ull << 32) : 0ull);
  }
  __device__ void emit(uint64_t i, uint64_t prefix) const
  {
    uint32_t node, bbox;
    locate(i, node, bbox);
    int const c = classify(t, bx, g, node, bbox);
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

__global__ void __launch_bounds__(256)
  gather_offsets_kernel(uint32_t const* __restrict__ out_node, uint32_t const* __restrict__ offset,
                        uint64_t k, uint32_t* __restrict__ keys)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    keys[i] = offset[out_node[i]];
}

__global__ void __launch_bounds__(256)
  permute_kernel(uint32_t const* __restrict__ perm, uint32_t const* __restrict__ out_node,
                 uint32_t const* __restrict__ out_bbox, uint64_t k, uint32_t* __restrict__ bbox_o,
                 uint32_t* __restrict__ quad_o)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t const p = perm[i];
    bbox_o[i]        = out_bbox[p];
    quad_o[i]        = out_node[p];
  }
}

// --- results back to job-local indices ---------------------------------------------
// Job-local (polygon, point) of hits [0, h), written at rows base + i. Hits are in
// job order (pairs follow the leaves' sorted positions), so a tile's first and
// last hits bound its jobs: per 4096-hit tile, waves 0 and 1 find those two jobs
// in parallel, and each hit's job is searched between them (usually equal). h is
// block-uniform, so every thread reaches every barrier.
__device__ __forceinline__ void local_hit_tiles(uint32_t const* __restrict__ poly,
                                                uint32_t const* __restrict__ pos, uint64_t h,
                                                uint64_t base, JobDev const* __restrict__ jobs,
                                                uint32_t num_jobs, uint32_t* __restrict__ out_poly,
                                                uint32_t* __restrict__ out_point)
{
  constexpr uint64_t kTile = 4096;
  __shared__ uint32_t s_jobs[2];
  int const w = threadIdx.x / kWave;
  for (uint64_t lo = static_cast<uint64_t>(blockIdx.x) * kTile; lo < h;
       lo += static_cast<uint64_t>(gridDim.x) * kTile) {
    uint64_t const hi = lo + kTile < h ? lo + kTile : h;
    __syncthreads();  // the previous tile's readers of s_jobs are done
    if (w < 2) {  // wave-uniform
      uint32_t const j = wave_find_job(jobs, 0, num_jobs - 1, pos[w == 0 ? lo : hi - 1]);
      if (lane_id() == 0) s_jobs[w] = j;
    }
    __syncthreads();
    uint32_t const j0 = s_jobs[0], j1 = s_jobs[1];
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      uint32_t const p = pos[i];
      uint32_t const j = j0 == j1 ? j0 : find_job(jobs, j0, j1, p);
      out_poly[base + i]  = poly[i] - jobs[j].poly_begin;
      out_point[base + i] = p - jobs[j].point_begin;
    }
  }
}

__global__ void __launch_bounds__(256)
  local_hits_kernel(uint32_t const* __restrict__ poly, uint32_t const* __restrict__ pos,
                    uint64_t h, JobDev const* __restrict__ jobs, uint32_t num_jobs,
                    uint32_t* __restrict__ out_poly, uint32_t* __restrict__ out_point)
{
  local_hit_tiles(poly, pos, h, 0, jobs, num_jobs, out_poly, out_point);
}

// First hit of every job (hits are in job order): first h with pos[h] >= point_begin.
__global__ void __launch_bounds__(256)
  job_hit_offsets_kernel(uint32_t const* __restrict__ pos, uint64_t h,
                         JobDev const* __restrict__ jobs, uint32_t num_jobs,
                         uint64_t* __restrict__ first)
{
  uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= num_jobs) return;
  uint32_t const b = jobs[j].point_begin;
  uint64_t lo = 0, hi = h;
  while (lo < hi) {
    uint64_t const mid = lo + ((hi - lo) >> 1);
    if (pos[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  first[j] = lo;
}

// --- bounding boxes ---------------------------------------------------------------------
// polygon_bounding_boxes of each job's own polygons (detail/bounding_boxes.cuh:139-184):
// one box per polygon with vertices, in order, then NaN rows for the job's empty
// polygons (reduce_by_key emits one box per non-empty geometry run).
__device__ __forceinline__ int64_t poly_first_vertex(int32_t const* part, int32_t const* ring,
                                                     int64_t num_ring_offsets, int64_t nv,
                                                     uint64_t p)
{
  int64_t const r = part[p];
  int64_t v       = (r >= 0 && r < num_ring_offsets) ? ring[r] : nv;
  if (v < 0) v = 0;
  if (v > nv) v = nv;
  return v;
}

struct NonEmptyOp {
  int32_t const* part;  // the group's polygon offsets (npg + 1)
  int32_t const* ring;
  int64_t num_ring_offsets;
  int64_t nv;
  uint32_t* prefix;  // npg + 1
  __device__ uint32_t count(uint64_t p) const
  {
    return poly_first_vertex(part, ring, num_ring_offsets, nv, p + 1) >
           poly_first_vertex(part, ring, num_ring_offsets, nv, p);
  }
  __device__ void emit(uint64_t p, uint32_t pre) const { prefix[p] = pre; }
};

// The box of polygon p, by one wave.
__device__ __forceinline__ void bbox_one(int32_t const* __restrict__ part,
                                         int32_t const* __restrict__ ring,
                                         int64_t num_ring_offsets, double const* __restrict__ vx,
                                         double const* __restrict__ vy, int64_t nv,
                                         uint32_t const* __restrict__ prefix, uint64_t p,
                                         JobDev const* __restrict__ jobs, uint32_t num_jobs,
                                         double* __restrict__ minx, double* __restrict__ miny,
                                         double* __restrict__ maxx, double* __restrict__ maxy)
{
  int64_t const v0 = poly_first_vertex(part, ring, num_ring_offsets, nv, p);
  int64_t const v1 = poly_first_vertex(part, ring, num_ring_offsets, nv, p + 1);
  if (v1 <= v0) return;
  // the job owning polygon p: last job whose poly_begin <= p (and that is not empty)
  uint32_t lo = 0, hi = num_jobs - 1;
  while (lo < hi) {
    uint32_t const mid = lo + ((hi - lo + 1) >> 1);
    if (jobs[mid].poly_begin <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  uint32_t const pb   = jobs[lo].poly_begin;
  uint64_t const dest = pb + (prefix[p] - prefix[pb]);
  double const nan    = __longlong_as_double(0x7FF8000000000000ll);
  double lx = nan, ly = nan, hx = nan, hy = nan;
  for (int64_t v = v0 + lane_id(); v < v1; v += kWave) {
    lx = box_min(lx, vx[v]);
    ly = box_min(ly, vy[v]);
    hx = box_max(hx, vx[v]);
    hy = box_max(hy, vy[v]);
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    lx = box_min(lx, __shfl_xor(lx, o, kWave));
    ly = box_min(ly, __shfl_xor(ly, o, kWave));
    hx = box_max(hx, __shfl_xor(hx, o, kWave));
    hy = box_max(hy, __shfl_xor(hy, o, kWave));
  }
  if (lane_id() == 0) {
    minx[dest] = lx;
    miny[dest] = ly;
    maxx[dest] = hx;
    maxy[dest] = hy;
  }
}

__global__ void __launch_bounds__(256)
  batch_bbox_kernel(int32_t const* __restrict__ part, int32_t const* __restrict__ ring,
                    int64_t num_ring_offsets, double const* __restrict__ vx,
                    double const* __restrict__ vy, int64_t nv, uint32_t const* __restrict__ prefix,
                    uint64_t npg, JobDev const* __restrict__ jobs, uint32_t num_jobs,
                    double* __restrict__ minx, double* __restrict__ miny, double* __restrict__ maxx,
                    double* __restrict__ maxy)
{
  uint64_t const waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
       p < npg; p += waves)
    bbox_one(part, ring, num_ring_offsets, vx, vy, nv, prefix, p, jobs, num_jobs, minx, miny, maxx,
             maxy);
}

__global__ void fill_nan_kernel(double* __restrict__ p, uint64_t n)
{
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = __longlong_as_double(0x7FF8000000000000ll);
}

// --- host -------------------------------------------------------------------------
struct HostJob {
  double min_x, max_x, min_y, max_y, scale;
  int d;
  int32_t max_size;
  int key_bits;
};

int bits_for(uint64_t v)  // bits needed for values 0..v
{
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

// Device descriptors of jobs [j0, j1), indices relative to the group's first
// point and polygon.
std::vector<JobDev> job_descs(std::vector<HostJob> const& hj, int64_t j0, int64_t j1,
                              int64_t const* jpo, int64_t const* jpoly)
{
  std::vector<JobDev> hd(static_cast<size_t>(j1 - j0));
  for (int64_t k = 0; k < j1 - j0; ++k) {
    HostJob const& h = hj[j0 + k];
    JobDev& d        = hd[k];
    d.min_x = h.min_x, d.min_y = h.min_y, d.max_x = h.max_x, d.max_y = h.max_y;
    d.scale       = h.scale;
    d.point_begin = static_cast<uint32_t>(jpo[j0 + k] - jpo[j0]);
    d.point_end   = static_cast<uint32_t>(jpo[j0 + k + 1] - jpo[j0]);
    d.poly_begin  = static_cast<uint32_t>(jpoly[j0 + k] - jpoly[j0]);
    d.poly_end    = static_cast<uint32_t>(jpoly[j0 + k + 1] - jpoly[j0]);
    d.outside_key = static_cast<uint32_t>((1u << (2 * h.d)) - 1u);
    d.d           = h.d;
    d.max_size    = static_cast<uint32_t>(h.max_size);
    d.shift0      = h.d >= 1 ? 2 * (h.d - 1) : 0;
  }
  return hd;
}

// Upper bound of the forest's node count for jobs [j0, j1): 64 level-0 nodes per
// job, then per level at most 4 children per internal node of the level above
// (at most points / (max_size + 1) of them) and at most 4^level nodes.
uint64_t forest_capacity(std::vector<HostJob> const& hj, int64_t j0, int64_t j1,
                         int64_t const* jpo)
{
  uint64_t cap = 64ull * static_cast<uint64_t>(j1 - j0);
  for (int64_t j = j0; j < j1; ++j) {
    uint64_t const nj = static_cast<uint64_t>(jpo[j + 1] - jpo[j]);
    uint64_t const ib = nj / (static_cast<uint64_t>(hj[j].max_size) + 1);
    for (int l = 1; l < std::max(hj[j].d, 1); ++l) {
      uint64_t c = std::min<uint64_t>(nj, 4 * ib);
      if (2 * l + 6 < 62) c = std::min<uint64_t>(c, uint64_t{64} << (2 * l));
      cap += c;
    }
  }
  QPIP_EXPECTS(cap < (uint64_t{1} << 32), "batched forest too large");
  return cap;
}

// Tile table of a group's point sort: each job's points are a segment of their own
// (the job order is the input order), sorted on the Morton key bits only; the last
// pass writes (job << key_bits) | Morton key, the forest's composite key.
void group_sort_tiles(int64_t j0, int64_t j1, int64_t const* jpo, int key_bits,
                      std::vector<SortTile>& tiles, std::vector<SortSegment>& segs)
{
  size_t const J = static_cast<size_t>(j1 - j0);
  std::vector<uint32_t> begin(J + 1), key_or(J);
  for (size_t k = 0; k <= J; ++k) begin[k] = static_cast<uint32_t>(jpo[j0 + k] - jpo[j0]);
  for (size_t k = 0; k < J; ++k)
    key_or[k] = key_bits < 32 ? static_cast<uint32_t>(k) << key_bits : 0u;
  make_sort_tiles(begin, tiles, segs, key_or);
}

// Groups of consecutive jobs whose job number and key fit 32 bits, < 2^31 points.
std::vector<std::pair<int64_t, int64_t>> job_groups(std::vector<HostJob> const& hj,
                                                    int64_t const* jpo)
{
  std::vector<std::pair<int64_t, int64_t>> out;
  int64_t const num_jobs = static_cast<int64_t>(hj.size());
  int64_t j0             = 0;
  while (j0 < num_jobs) {
    int64_t j1 = j0 + 1;
    int kb     = hj[j0].key_bits;
    while (j1 < num_jobs) {
      int const kb2 = std::max(kb, hj[j1].key_bits);
      if (kb2 + bits_for(static_cast<uint64_t>(j1 - j0)) > 32) break;
      if (jpo[j1 + 1] - jpo[j0] >= (int64_t{1} << 31)) break;
      kb = kb2;
      ++j1;
    }
    QPIP_EXPECTS(jpo[j1] - jpo[j0] < (int64_t{1} << 31), "a job must have fewer than 2^31 points");
    out.emplace_back(j0, j1);
    j0 = j1;
  }
  return out;
}

// Validates the job descriptions and applies the reference's clamps
// (indexing.py:165-189, detail/point_quadtree.cuh:259-268,
// quadtree_bbox_filtering.cu:102-106 for jobs that reach the join).
std::vector<HostJob> prepare_jobs(int64_t num_jobs, int64_t const* job_point_offsets,
                                  int64_t const* job_polygon_offsets, double const* job_bbox,
                                  double const* job_scale, int8_t const* job_max_depth,
                                  int32_t const* job_max_size, int64_t num_points,
                                  int64_t num_poly_offsets, int64_t num_poly_x,
                                  int64_t num_poly_y)
{
  QPIP_EXPECTS(job_point_offsets && job_polygon_offsets && job_bbox && job_scale &&
                 job_max_depth && job_max_size,
               "job descriptions must not be null");
  QPIP_EXPECTS(num_poly_x == num_poly_y,
               "numbers of vertices must be the same for both x and y columns");
  QPIP_EXPECTS(job_point_offsets[0] == 0 && job_point_offsets[num_jobs] == num_points,
               "job point offsets must start at 0 and end at the number of points");
  int64_t const num_polys = num_poly_offsets > 0 ? num_poly_offsets - 1 : 0;
  QPIP_EXPECTS(job_polygon_offsets[0] == 0 && job_polygon_offsets[num_jobs] == num_polys,
               "job polygon offsets must start at 0 and end at the number of polygons");
  std::vector<HostJob> hj(num_jobs);
  for (int64_t j = 0; j < num_jobs; ++j) {
    QPIP_EXPECTS(job_point_offsets[j + 1] >= job_point_offsets[j] &&
                   job_polygon_offsets[j + 1] >= job_polygon_offsets[j],
                 "job offsets must not decrease");
    HostJob& h      = hj[j];
    double const x0 = job_bbox[4 * j], x1 = job_bbox[4 * j + 1];
    double const y0 = job_bbox[4 * j + 2], y1 = job_bbox[4 * j + 3];
    h.min_x    = std::min(x0, x1);
    h.max_x    = std::max(x0, x1);
    h.min_y    = std::min(y0, y1);
    h.max_y    = std::max(y0, y1);
    h.max_size = std::max(1, job_max_size[j]);
    h.d        = std::max(0, std::min(15, static_cast<int>(job_max_depth[j])));
    h.scale    = std::max(job_scale[j],
                       std::max(h.max_x - h.min_x, h.max_y - h.min_y) / ((1 << h.d) + 2));
    if (job_point_offsets[j + 1] > job_point_offsets[j] &&
        job_polygon_offsets[j + 1] > job_polygon_offsets[j]) {
      QPIP_EXPECTS(h.scale > 0, "scale must be positive");
      QPIP_EXPECTS(h.min_x < h.max_x && h.min_y < h.max_y,
                   "invalid bounding box (x_min, x_max, y_min, y_max)");
      QPIP_EXPECTS(h.d > 0 && h.d < 16, "maximum depth must be positive and less than 16");
    }
    int cbits = 0;
    while ((uint64_t{1} << cbits) <= (uint64_t{1} << h.d) + 2) ++cbits;
    h.key_bits = std::min(32, 2 * cbits);
  }
  return hj;
}

struct GroupResult {
  dbuf<uint32_t> poly, point;  // job-local hits of the group
  uint64_t hits = 0;
};

// Runs one group of jobs [j0, j1). Writes point_indices for the group's points into
// out_pi (already offset) and returns the group's hits and per-job hit counts.
void run_group(std::vector<HostJob> const& hj, int64_t j0, int64_t j1, int64_t const* jpo,
               int64_t const* jpoly, double const* points_xy, int32_t const* poly_offsets,
               int32_t const* ring_offsets, int64_t num_ring_offsets, double const* poly_x,
               double const* poly_y, int64_t num_vertices, uint32_t* out_pi, GroupResult& res,
               std::vector<uint64_t>& job_hits, hipStream_t stream)
{
  uint32_t const J   = static_cast<uint32_t>(j1 - j0);
  int64_t const P0   = jpo[j0], P1 = jpo[j1];
  int64_t const PO0  = jpoly[j0], PO1 = jpoly[j1];
  uint64_t const n   = static_cast<uint64_t>(P1 - P0);
  uint64_t const npg = static_cast<uint64_t>(PO1 - PO0);
  int key_bits = 1, maxd = 0;
  for (int64_t j = j0; j < j1; ++j) {
    key_bits = std::max(key_bits, hj[j].key_bits);
    maxd     = std::max(maxd, hj[j].d);
  }

  std::vector<JobDev> hd = job_descs(hj, j0, j1, jpo, jpoly);
  std::vector<SortTile> tile_tab;
  std::vector<SortSegment> seg_tab;
  group_sort_tiles(j0, j1, jpo, key_bits, tile_tab, seg_tab);
  dbuf<JobDev> jobs(J);
  dbuf<SortTile> stiles(tile_tab.size());
  dbuf<SortSegment> ssegs(seg_tab.size());
  QPIP_HIP(hipMemcpyAsync(jobs.get(), hd.data(), J * sizeof(JobDev), hipMemcpyHostToDevice, stream));
  if (!tile_tab.empty())
    QPIP_HIP(hipMemcpyAsync(stiles.get(), tile_tab.data(), tile_tab.size() * sizeof(SortTile),
                            hipMemcpyHostToDevice, stream));
  QPIP_HIP(hipMemcpyAsync(ssegs.get(), seg_tab.data(), seg_tab.size() * sizeof(SortSegment),
                          hipMemcpyHostToDevice, stream));
  QPIP_HIP(hipStreamSynchronize(stream));  // hd, tile_tab, seg_tab are pageable host memory
  for (uint32_t k = 0; k < J; ++k) job_hits[j0 + k] = 0;
  if (n == 0) return;

  // 1-2. keys and the sort, each job's points on their own; the last pass writes
  // the job-local point_indices output (the PIP test adds each node's job base)
  dbuf<uint32_t> keys_a(n), keys_b(n), vals_tmp(n), samples((n + 31) / 32);
  dbuf<uint32_t> counts0(tile_tab.size() * kRadix);
  {
    KernelTimer kt("batch_keygen", stream);
    batch_keygen_seg_kernel<<<static_cast<unsigned>(tile_tab.size()), 256, 0, stream>>>(
      points_xy + 2 * P0, stiles.get(), jobs.get(), keys_a.get(), counts0.get());
    QPIP_LAUNCH_CHECK();
  }
  SortResult sorted = segmented_radix_sort_pairs(keys_a.get(), keys_b.get(), out_pi,
                                                 vals_tmp.get(), n, key_bits, stiles.get(),
                                                 tile_tab.size(), ssegs.get(), seg_tab.size(),
                                                 counts0.get(), nullptr, samples.get(), true,
                                                 stream, "batch_sort");
  vals_tmp.release();

  // 3. forest levels
  uint64_t const cap = forest_capacity(hj, j0, j1, jpo);
  dbuf<uint32_t> nkey(cap), nstart(cap), ncount(cap), cfirst(cap), ccount(cap), njob(cap);
  dbuf<uint8_t> nlevel(cap);
  dbuf<uint4> split(cap), staging(static_cast<size_t>(J) * 64);
  dbuf<JobTree> trees(J);
  dbuf<ForestMeta> fmeta(1);
  dbuf<uint32_t> partials(std::max(kLevelBlocks, kScanBlocks) + 1);
  QPIP_HIP(hipMemsetAsync(fmeta.get(), 0, sizeof(ForestMeta), stream));
  {
    KernelTimer kt("batch_levels", stream);
    batch_level0_kernel<<<(J + 3) / 4, 256, 0, stream>>>(sorted.keys, jobs.get(), J, key_bits,
                                                         staging.get(), trees.get());
    QPIP_LAUNCH_CHECK();
    Level0Op l0{trees.get(), staging.get(), nkey.get(), nstart.get(), ncount.get(), nlevel.get(),
                njob.get()};
    count_scan_emit<Level0Op, uint32_t, uint64_t>(l0, nullptr, J, partials.get(),
                                                  &fmeta.get()->sizes[0], stream,
                                                  scan_blocks_for(J), "batch_level0");
    JobLevels rule{njob.get(), jobs.get(), key_bits};
    for (int l = 0; l + 2 <= maxd; ++l) {
      LevelArgs la{sorted.keys, nkey.get(), nstart.get(), ncount.get(), nlevel.get(),
                   cfirst.get(), ccount.get(), split.get(), partials.get(), fmeta.get()->sizes,
                   &fmeta.get()->overflow, cap, l, samples.get()};
      level_split_kernel<<<kLevelBlocks, 256, 0, stream>>>(la, rule);
      QPIP_LAUNCH_CHECK();
      level_emit_kernel<<<kLevelBlocks, 256, 0, stream>>>(la, rule);
      QPIP_LAUNCH_CHECK();
    }
  }
  ForestMeta hm;
  QPIP_HIP(hipMemcpyAsync(&hm, fmeta.get(), sizeof(ForestMeta), hipMemcpyDeviceToHost, stream));
  QPIP_HIP(hipStreamSynchronize(stream));
  if (hm.overflow) throw device_error("batched forest node capacity exceeded (internal error)");
  uint64_t M = 0;
  for (int l = 0; l < std::max(maxd, 1); ++l) M += hm.sizes[l];
  batch_marked_kernel<<<(J + 255) / 256, 256, 0, stream>>>(nkey.get(), nstart.get(), ncount.get(),
                                                          njob.get(), jobs.get(), J, fmeta.get(),
                                                          trees.get());
  QPIP_LAUNCH_CHECK();
  dbuf<uint8_t> f_internal(std::max<uint64_t>(M, 1));
  dbuf<uint32_t> f_length(std::max<uint64_t>(M, 1)), f_offset(std::max<uint64_t>(M, 1)),
    f_base(std::max<uint64_t>(M, 1));
  if (M > 0) {
    batch_finalize_kernel<<<grid_for(M, 256), 256, 0, stream>>>(
      nkey.get(), nstart.get(), ncount.get(), nlevel.get(), njob.get(), cfirst.get(), ccount.get(),
      jobs.get(), trees.get(), M, nullptr, 0, f_internal.get(), f_length.get(), f_offset.get(),
      f_base.get());
    QPIP_LAUNCH_CHECK();
  }
  if (M == 0 || npg == 0) {
    QPIP_HIP(hipStreamSynchronize(stream));
    return;
  }

  // 4. bounding boxes of each job's polygons
  dbuf<double> boxes(4 * npg);
  double* bminx = boxes.get();
  double* bminy = bminx + npg;
  double* bmaxx = bminy + npg;
  double* bmaxy = bmaxx + npg;
  {
    KernelTimer kt("batch_bbox", stream);
    fill_nan_kernel<<<grid_for(4 * npg, 256), 256, 0, stream>>>(bminx, 4 * npg);
    QPIP_LAUNCH_CHECK();
    dbuf<uint32_t> prefix(npg + 1);
    NonEmptyOp ne{poly_offsets + PO0, ring_offsets, num_ring_offsets, num_vertices, prefix.get()};
    count_scan_emit<NonEmptyOp, uint32_t, uint32_t>(ne, nullptr, npg, partials.get(),
                                                    prefix.get() + npg, stream,
                                                    scan_blocks_for(npg), "batch_bbox_scan");
    batch_bbox_kernel<<<grid_for(npg, 256 / kWave), 256, 0, stream>>>(
      poly_offsets + PO0, ring_offsets, num_ring_offsets, poly_x, poly_y, num_vertices,
      prefix.get(), npg, jobs.get(), J, bminx, bminy, bmaxx, bmaxy);
    QPIP_LAUNCH_CHECK();
  }
  int64_t const nb = static_cast<int64_t>(npg);

  // 5. join
  TreeView tv{nkey.get(), nlevel.get(), f_internal.get(), f_length.get(), f_offset.get(), M};
  BoxView bv{bminx, bminy, bmaxx, bmaxy, static_cast<uint64_t>(nb)};
  JobGeometry geom{njob.get(), jobs.get()};
  dbuf<uint64_t> top_off(J);
  dbuf<JoinMeta> jmeta(1);
  dbuf<uint64_t> partials64(kScanBlocks + 1);
  uint64_t cap_out = std::max<uint64_t>(uint64_t{1} << 20, 8 * npg);
  uint64_t cap_fr  = cap_out;
  JoinMeta h_meta{};
  dbuf<uint32_t> out_node, out_bbox, fa_node, fa_bbox, fb_node, fb_bbox;
  for (int attempt = 0;; ++attempt) {
    out_node.reset(cap_out);
    out_bbox.reset(cap_out);
    fa_node.reset(cap_fr);
    fa_bbox.reset(cap_fr);
    fb_node.reset(cap_fr);
    fb_bbox.reset(cap_fr);
    QPIP_HIP(hipMemsetAsync(jmeta.get(), 0, sizeof(JoinMeta), stream));
    JoinMeta* dm = jmeta.get();
    TopSizeOp ts{trees.get(), jobs.get(), top_off.get()};
    count_scan_emit<TopSizeOp, uint64_t, uint64_t>(ts, nullptr, J, partials64.get(),
                                                   &dm->top_pairs, stream, scan_blocks_for(J),
                                                   "batch_join_setup");
    BatchTopOp top{tv, bv, geom, trees.get(), jobs.get(), top_off.get(), J, dm, out_node.get(),
                   out_bbox.get(), fa_node.get(), fa_bbox.get(), cap_out, cap_fr};
    count_scan_emit<BatchTopOp, uint64_t, SplitTotal>(top, &dm->top_pairs, 0, partials64.get(),
                                                      &dm->level[0], stream, kScanBlocks,
                                                      "batch_join_level");
    uint32_t* cur_node = fa_node.get();
    uint32_t* cur_bbox = fa_bbox.get();
    uint32_t* nxt_node = fb_node.get();
    uint32_t* nxt_bbox = fb_bbox.get();
    for (int lvl = 1; lvl < maxd; ++lvl) {
      DescendOp<JobGeometry> op{tv, bv, geom, dm, cur_node, cur_bbox, out_node.get(),
                                out_bbox.get(), nxt_node, nxt_bbox, cap_out, cap_fr, lvl};
      count_scan_emit<DescendOp<JobGeometry>, uint64_t, SplitTotal>(
        op, &dm->level[lvl - 1].quad, 0, partials64.get(), &dm->level[lvl], stream, kScanBlocks,
        "batch_join_level");
      std::swap(cur_node, nxt_node);
      std::swap(cur_bbox, nxt_bbox);
    }
    QPIP_HIP(hipMemcpyAsync(&h_meta, dm, sizeof(JoinMeta), hipMemcpyDeviceToHost, stream));
    QPIP_HIP(hipStreamSynchronize(stream));
    if (h_meta.bad_index) throw logic_error("cuSpatial failure: forest child range out of range");
    if (!h_meta.overflow) break;
    uint64_t tot_out = 0, max_fr = h_meta.top_pairs;
    for (int l = 0; l < std::max(maxd, 1); ++l) {
      tot_out += h_meta.level[l].leaf;
      max_fr = std::max<uint64_t>(max_fr, h_meta.level[l].quad);
    }
    cap_out = std::max<uint64_t>(cap_out * 2, tot_out + tot_out / 4 + 1024);
    cap_fr  = std::max<uint64_t>(cap_fr * 2, max_fr + max_fr / 4 + 1024);
    if (attempt > 8) throw device_error("join capacity growth did not converge");
  }
  uint64_t K = 0;
  for (int l = 0; l < std::max(maxd, 1); ++l) K += h_meta.level[l].leaf;
  if (K == 0) return;
  QPIP_EXPECTS(K < (uint64_t{1} << 32), "join produced more than 2^32 pairs");
  dbuf<uint32_t> ka(K), kb(K), va(K), vb(K), pb(K), pq(K);
  gather_offsets_kernel<<<grid_for(K, 256), 256, 0, stream>>>(out_node.get(), f_offset.get(), K,
                                                              ka.get());
  QPIP_LAUNCH_CHECK();
  SortResult ps = radix_sort_pairs(ka.get(), kb.get(), nullptr, va.get(), vb.get(), K,
                                   std::max(1, bits_for(n)), stream, nullptr, "batch_join_sort");
  permute_kernel<<<grid_for(K, 256), 256, 0, stream>>>(ps.values, out_node.get(), out_bbox.get(),
                                                       K, pb.get(), pq.get());
  QPIP_LAUNCH_CHECK();

  // 6. point-in-polygon on the forest
  uint32_t *hp = nullptr, *ht = nullptr;
  int64_t H = 0;
  int rc = quadtree_point_in_polygon_entry(
    pb.get(), pq.get(), static_cast<int64_t>(K), nullptr, nullptr, nullptr, f_length.get(),
    f_offset.get(), static_cast<int64_t>(M), out_pi, static_cast<int64_t>(n), nullptr, nullptr,
    points_xy + 2 * P0, static_cast<int64_t>(n), static_cast<int64_t>(n),
    reinterpret_cast<uint32_t const*>(poly_offsets + PO0), static_cast<int64_t>(npg) + 1,
    reinterpret_cast<uint32_t const*>(ring_offsets), num_ring_offsets, poly_x, poly_y,
    num_vertices, num_vertices, nullptr, stream, &hp, &ht, &H, f_base.get());
  if (rc != QPIP_SUCCESS) throw logic_error(qpip_last_error_message());
  struct Hits {
    uint32_t* p[2];
    ~Hits()
    {
      qpip_free(p[0]);
      qpip_free(p[1]);
    }
  } hits{{hp, ht}};
  if (H == 0) return;
  uint64_t const h = static_cast<uint64_t>(H);
  res.poly.reset(h);
  res.point.reset(h);
  res.hits = h;
  // one block per 4096-hit tile
  local_hits_kernel<<<grid_for(h, 4096), 256, 0, stream>>>(hp, ht, h, jobs.get(), J,
                                                           res.poly.get(), res.point.get());
  QPIP_LAUNCH_CHECK();
  dbuf<uint64_t> first(J);
  job_hit_offsets_kernel<<<(J + 255) / 256, 256, 0, stream>>>(ht, h, jobs.get(), J, first.get());
  QPIP_LAUNCH_CHECK();
  std::vector<uint64_t> hf(J);
  QPIP_HIP(hipMemcpyAsync(hf.data(), first.get(), J * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          stream));
  QPIP_HIP(hipStreamSynchronize(stream));
  for (uint32_t k = 0; k < J; ++k) job_hits[j0 + k] = (k + 1 < J ? hf[k + 1] : h) - hf[k];
}

// --- captured plan: the same pipeline with every size on the device ----------------
// A plan fixes the jobs, the polygons and the point buffer's address; each launch
// replays one HIP graph of the whole pipeline over whatever points the buffer then
// holds. The polygon boxes and edge table depend on nothing else and are built once
// at creation. Data-dependent sizes (forest nodes, leaf pairs, PIP chunks, hits)
// stay on the device: every kernel after them runs on a grid sized by a capacity
// and reads the actual count, every write is bounded by its capacity, and a count
// past its capacity sets a flag that the result call reports. Capacities come from
// sizing runs over the points in the buffer at creation, times a headroom factor.

// Counts of a group whose join is not run (no points or no polygons): no hits.
__global__ void plan_empty_kernel(DeviceCounts* cnt, DeviceCounts const* prev,
                                  ForestMeta const* fm, int64_t* job_first, uint32_t num_jobs)
{
  uint32_t const j    = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t const base = prev ? prev->hit_base + prev->hits : 0;
  if (j == 0) {
    cnt->pairs    = 0;
    cnt->chunks   = 0;
    cnt->hits     = 0;
    cnt->hit_base = base;
    cnt->flags    = (fm && fm->overflow) ? 2u : 0u;
  }
  if (j < num_jobs) job_first[j] = static_cast<int64_t>(base);
}

// Leaf pairs of the join, clamped to the pair capacity; starts the group's flags.
__global__ void plan_pairs_kernel(JoinMeta const* jm, int levels, uint64_t cap_out,
                                  ForestMeta const* fm, DeviceCounts* cnt,
                                  DeviceCounts const* prev)
{
  if (threadIdx.x != 0) return;
  uint64_t k = 0;
  for (int l = 0; l < levels; ++l) k += jm->level[l].leaf;
  uint32_t f = 0;
  if (jm->bad_index) f |= 1u;
  if (jm->overflow || fm->overflow || k > cap_out) f |= 2u;
  cnt->pairs    = k < cap_out ? k : cap_out;
  cnt->chunks   = 0;
  cnt->hits     = 0;
  cnt->hit_base = prev ? prev->hit_base + prev->hits : 0;
  cnt->flags    = f;
}

// Sort keys of the leaf pairs (leaf offsets); rows past the pair count get the
// largest key, so the stable sort leaves them after every real pair.
__global__ void __launch_bounds__(256)
  plan_gather_kernel(uint32_t const* __restrict__ out_node, uint32_t const* __restrict__ offset,
                     DeviceCounts const* cnt, uint64_t cap, uint32_t* __restrict__ keys)
{
  uint64_t const k = cnt->pairs;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < cap;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    keys[i] = i < k ? offset[out_node[i]] : 0xFFFFFFFFu;
}

__global__ void __launch_bounds__(256)
  plan_permute_kernel(uint32_t const* __restrict__ perm, uint32_t const* __restrict__ out_node,
                      uint32_t const* __restrict__ out_bbox, DeviceCounts const* cnt,
                      uint32_t* __restrict__ bbox_o, uint32_t* __restrict__ quad_o)
{
  uint64_t const k = cnt->pairs;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t const p = perm[i];
    bbox_o[i]        = out_bbox[p];
    quad_o[i]        = out_node[p];
  }
}

// Job-local hits, written at the group's first output row.
__global__ void __launch_bounds__(256)
  plan_local_hits_kernel(uint32_t const* __restrict__ poly, uint32_t const* __restrict__ pos,
                         DeviceCounts* cnt, JobDev const* __restrict__ jobs, uint32_t num_jobs,
                         uint64_t out_cap, uint32_t* __restrict__ out_poly,
                         uint32_t* __restrict__ out_point)
{
  uint64_t const base = cnt->hit_base;
  uint64_t h          = cnt->hits;
  // rows past out_cap are dropped and flagged; the count stays block-uniform
  uint64_t const room = out_cap > base ? out_cap - base : 0;
  if (h > room) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&cnt->flags, 2u);
    h = room;
  }
  local_hit_tiles(poly, pos, h, base, jobs, num_jobs, out_poly, out_point);
}

// First output row of every job of the group.
__global__ void __launch_bounds__(256)
  plan_job_first_kernel(uint32_t const* __restrict__ pos, DeviceCounts const* cnt,
                        JobDev const* __restrict__ jobs, uint32_t num_jobs,
                        int64_t* __restrict__ job_first)
{
  uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= num_jobs) return;
  uint32_t const b = jobs[j].point_begin;
  uint64_t lo = 0, hi = cnt->hits;
  while (lo < hi) {
    uint64_t const mid = lo + ((hi - lo) >> 1);
    if (pos[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  job_first[j] = static_cast<int64_t>(cnt->hit_base + lo);
}

template <class T>
void zero_struct(T* p, hipStream_t stream)
{
  static_assert(sizeof(T) % 4 == 0, "cleared as 32-bit words");
  fill_u32(reinterpret_cast<uint32_t*>(p), sizeof(T) / 4, 0u, stream);
}

struct PlanCaps {
  uint64_t pairs = 0, frontier = 0, chunks = 0, hits = 0;
};

// One group's fixed description and every buffer its launch sequence touches.
struct PlanGroup {
  int64_t j0 = 0, j1 = 0;
  uint32_t J = 0;
  int64_t P0 = 0, PO0 = 0;
  uint64_t n = 0, npg = 0;
  int key_bits = 1, maxd = 0;
  uint64_t node_cap = 0;
  bool tree = false, join = false;  // stages that run (host-known)
  PlanCaps cap;
  dbuf<JobDev> jobs;
  // tree
  dbuf<uint32_t> keys_a, keys_b, vals_tmp, counts0, samples;
  dbuf<SortTile> stiles;  // the segmented point sort's tile table
  dbuf<SortSegment> ssegs;
  uint64_t num_stiles = 0;
  dbuf<uint32_t> nkey, nstart, ncount, cfirst, ccount, njob, partials;
  dbuf<uint8_t> nlevel, f_internal;
  dbuf<uint32_t> f_length, f_offset, f_base;
  dbuf<uint4> split, staging;
  dbuf<JobTree> trees;
  dbuf<ForestMeta> fmeta;
  // boxes (built at creation)
  dbuf<double> boxes;
  // join and pair sort (capacity-sized)
  dbuf<uint64_t> top_off, partials64;
  dbuf<JoinMeta> jmeta;
  dbuf<uint32_t> out_node, out_bbox, fa_node, fa_bbox, fb_node, fb_bbox;
  dbuf<uint32_t> ka, kb, va, vb, pb, pq, pair_counts, pair_sums;
  // point-in-polygon (capacity-sized)
  dbuf<uint32_t> chunk_start, hit_poly, hit_pos;
  dbuf<uint64_t> pip_partials, pip_totals, masks, hit_start;
  // the polygon-major PIP test's pair order (used when pip_polygon_group > 0)
  dbuf<uint32_t> pm_count, pm_start, pm_cursor;
  dbuf<uint4> pm_order, pm_units;
  dbuf<uint64_t> pm_partials, pm_total, pm_num_units;
  dbuf<uint2> chunk_info;
  uint32_t pairs_per_wave = 16;
};

}  // namespace

struct BatchPlan {
  int device = 0;
  hipStream_t cstream = nullptr;  // capture stream
  std::vector<HostJob> hj;
  std::vector<int64_t> jpo, jpoly;
  int64_t num_jobs = 0, num_points = 0;
  double const* points_xy = nullptr;
  int32_t const* poly_offsets = nullptr;
  int32_t const* ring_offsets = nullptr;
  int64_t num_ring_offsets = 0, num_vertices = 0;
  double const *poly_x = nullptr, *poly_y = nullptr;
  dbuf<uint32_t> ring_edges;
  dbuf<LineEdge> lines;    // per edge: the PIP test's line form
  dbuf<double> ring_amax;  // per ring
  dbuf<PolyBox> poly_box;  // per polygon of the batch
  std::vector<std::unique_ptr<PlanGroup>> groups;
  dbuf<DeviceCounts> counts;  // one per group
  dbuf<int64_t> job_first;    // per job: first output row
  dbuf<uint32_t> out_pi, out_poly, out_point;
  uint64_t hit_cap = 0;
  hipGraph_t graph    = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t last_stream = nullptr;
  bool launched = false;

  ~BatchPlan()
  {
    if (launched) (void)hipStreamSynchronize(last_stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cstream) {
      (void)hipStreamSynchronize(cstream);
      (void)hipStreamDestroy(cstream);
    }
    // every stream that used the buffers has finished: they go back idle
    CompletedScope done;
    groups.clear();
    ring_edges.release();
    lines.release();
    ring_amax.release();
    poly_box.release();
    counts.release();
    job_first.release();
    out_pi.release();
    out_poly.release();
    out_point.release();
  }
};

namespace {

// Allocates the capacity-sized buffers of a group (join, pair sort, PIP).
void plan_alloc(PlanGroup& g)
{
  PlanCaps const& c = g.cap;
  g.out_node.reset(c.pairs);
  g.out_bbox.reset(c.pairs);
  g.fa_node.reset(c.frontier);
  g.fa_bbox.reset(c.frontier);
  g.fb_node.reset(c.frontier);
  g.fb_bbox.reset(c.frontier);
  g.ka.reset(c.pairs);
  g.kb.reset(c.pairs);
  g.va.reset(c.pairs);
  g.vb.reset(c.pairs);
  g.pb.reset(c.pairs);
  g.pq.reset(c.pairs);
  g.pair_counts.reset(sort_counts_elems(c.pairs));
  g.pair_sums.reset(sort_sums_elems(c.pairs));
  g.chunk_start.reset(c.pairs + 1);
  g.masks.reset(c.chunks);
  g.chunk_info.reset(c.chunks);
  g.pm_order.reset(c.pairs);
  g.pm_units.reset(c.pairs);
  g.hit_start.reset(c.chunks);
  g.hit_poly.reset(c.hits);
  g.hit_pos.reset(c.hits);
}

// Enqueues group g's launch sequence: launches only (no allocation, no host
// synchronisation, no memset node), so that it can be captured.
void plan_enqueue(BatchPlan& P, PlanGroup& g, size_t gi, hipStream_t s)
{
  DeviceCounts* cnt        = P.counts.get() + gi;
  DeviceCounts const* prev = gi > 0 ? P.counts.get() + gi - 1 : nullptr;
  int64_t* job_first       = P.job_first.get() + g.j0;
  if (!g.tree) {
    plan_empty_kernel<<<grid_for(g.J, 256), 256, 0, s>>>(cnt, prev, nullptr, job_first, g.J);
    QPIP_LAUNCH_CHECK();
    return;
  }
  double const* xy = P.points_xy + 2 * g.P0;
  // 1-2. keys and the sort, each job's points on their own
  {
    KernelTimer kt("batch_keygen", s);
    batch_keygen_seg_kernel<<<static_cast<unsigned>(g.num_stiles), 256, 0, s>>>(
      xy, g.stiles.get(), g.jobs.get(), g.keys_a.get(), g.counts0.get());
    QPIP_LAUNCH_CHECK();
  }
  // the last pass writes the job-local point_indices output, which the PIP test
  // reads with each node's job base (point_base)
  SortResult sorted = segmented_radix_sort_pairs(
    g.keys_a.get(), g.keys_b.get(), P.out_pi.get() + g.P0, g.vals_tmp.get(), g.n, g.key_bits,
    g.stiles.get(), g.num_stiles, g.ssegs.get(), g.J, g.counts0.get(), nullptr, g.samples.get(),
    true, s, "batch_sort");
  // 3. forest levels
  zero_struct(g.fmeta.get(), s);
  {
    KernelTimer kt("batch_levels", s);
    batch_level0_kernel<<<(g.J + 3) / 4, 256, 0, s>>>(sorted.keys, g.jobs.get(), g.J, g.key_bits,
                                                      g.staging.get(), g.trees.get());
    QPIP_LAUNCH_CHECK();
    Level0Op l0{g.trees.get(), g.staging.get(), g.nkey.get(), g.nstart.get(), g.ncount.get(),
                g.nlevel.get(), g.njob.get()};
    count_scan_emit<Level0Op, uint32_t, uint64_t>(l0, nullptr, g.J, g.partials.get(),
                                                  &g.fmeta.get()->sizes[0], s,
                                                  scan_blocks_for(g.J), "batch_level0");
    JobLevels rule{g.njob.get(), g.jobs.get(), g.key_bits};
    for (int l = 0; l + 2 <= g.maxd; ++l) {
      LevelArgs la{sorted.keys, g.nkey.get(), g.nstart.get(), g.ncount.get(), g.nlevel.get(),
                   g.cfirst.get(), g.ccount.get(), g.split.get(), g.partials.get(),
                   g.fmeta.get()->sizes, &g.fmeta.get()->overflow, g.node_cap, l,
                   g.samples.get()};
      level_split_kernel<<<kLevelBlocks, 256, 0, s>>>(la, rule);
      QPIP_LAUNCH_CHECK();
      level_emit_kernel<<<kLevelBlocks, 256, 0, s>>>(la, rule);
      QPIP_LAUNCH_CHECK();
    }
  }
  int const levels = std::max(g.maxd, 1);
  batch_marked_kernel<<<(g.J + 255) / 256, 256, 0, s>>>(g.nkey.get(), g.nstart.get(),
                                                        g.ncount.get(), g.njob.get(),
                                                        g.jobs.get(), g.J, g.fmeta.get(),
                                                        g.trees.get());
  QPIP_LAUNCH_CHECK();
  // a bounded grid striding over the node count (the capacity is far above it)
  batch_finalize_kernel<<<grid_for(g.node_cap, 256, 8192), 256, 0, s>>>(
    g.nkey.get(), g.nstart.get(), g.ncount.get(), g.nlevel.get(), g.njob.get(), g.cfirst.get(),
    g.ccount.get(), g.jobs.get(), g.trees.get(), 0, g.fmeta.get(), levels, g.f_internal.get(),
    g.f_length.get(), g.f_offset.get(), g.f_base.get());
  QPIP_LAUNCH_CHECK();
  if (!g.join) {
    plan_empty_kernel<<<grid_for(g.J, 256), 256, 0, s>>>(cnt, prev, g.fmeta.get(), job_first,
                                                         g.J);
    QPIP_LAUNCH_CHECK();
    return;
  }
  // 5. join over the forest (node arrays read up to the node capacity)
  uint64_t const npg = g.npg;
  double* const bminx = g.boxes.get();
  TreeView tv{g.nkey.get(), g.nlevel.get(), g.f_internal.get(), g.f_length.get(),
              g.f_offset.get(), g.node_cap};
  BoxView bv{bminx, bminx + npg, bminx + 2 * npg, bminx + 3 * npg, npg};
  JobGeometry geom{g.njob.get(), g.jobs.get()};
  zero_struct(g.jmeta.get(), s);
  JoinMeta* dm = g.jmeta.get();
  TopSizeOp ts{g.trees.get(), g.jobs.get(), g.top_off.get()};
  count_scan_emit<TopSizeOp, uint64_t, uint64_t>(ts, nullptr, g.J, g.partials64.get(),
                                                 &dm->top_pairs, s, scan_blocks_for(g.J),
                                                 "batch_join_setup");
  BatchTopOp top{tv, bv, geom, g.trees.get(), g.jobs.get(), g.top_off.get(), g.J, dm,
                 g.out_node.get(), g.out_bbox.get(), g.fa_node.get(), g.fa_bbox.get(),
                 g.cap.pairs, g.cap.frontier};
  count_scan_emit<BatchTopOp, uint64_t, SplitTotal>(top, &dm->top_pairs, 0, g.partials64.get(),
                                                    &dm->level[0], s, kScanBlocks,
                                                    "batch_join_level");
  uint32_t* cur_node = g.fa_node.get();
  uint32_t* cur_bbox = g.fa_bbox.get();
  uint32_t* nxt_node = g.fb_node.get();
  uint32_t* nxt_bbox = g.fb_bbox.get();
  for (int lvl = 1; lvl < g.maxd; ++lvl) {
    DescendOp<JobGeometry> op{tv, bv, geom, dm, cur_node, cur_bbox, g.out_node.get(),
                              g.out_bbox.get(), nxt_node, nxt_bbox, g.cap.pairs, g.cap.frontier,
                              lvl};
    count_scan_emit<DescendOp<JobGeometry>, uint64_t, SplitTotal>(
      op, &dm->level[lvl - 1].quad, 0, g.partials64.get(), &dm->level[lvl], s, kScanBlocks,
      "batch_join_level");
    std::swap(cur_node, nxt_node);
    std::swap(cur_bbox, nxt_bbox);
  }
  plan_pairs_kernel<<<1, 64, 0, s>>>(dm, levels, g.cap.pairs, g.fmeta.get(), cnt, prev);
  QPIP_LAUNCH_CHECK();
  plan_gather_kernel<<<grid_for(g.cap.pairs, 256), 256, 0, s>>>(
    g.out_node.get(), g.f_offset.get(), cnt, g.cap.pairs, g.ka.get());
  QPIP_LAUNCH_CHECK();
  SortResult ps = radix_sort_pairs(g.ka.get(), g.kb.get(), nullptr, g.va.get(), g.vb.get(),
                                   g.cap.pairs, std::max(1, bits_for(g.n)), s, nullptr,
                                   "batch_join_sort", g.pair_counts.get(), g.pair_sums.get());
  plan_permute_kernel<<<grid_for(g.cap.pairs, 256), 256, 0, s>>>(
    ps.values, g.out_node.get(), g.out_bbox.get(), cnt, g.pb.get(), g.pq.get());
  QPIP_LAUNCH_CHECK();
  // 6. point-in-polygon on the forest
  PipDevice d{};
  d.pair_poly      = g.pb.get();
  d.pair_quad      = g.pq.get();
  d.pair_cap       = g.cap.pairs;
  d.length         = g.f_length.get();
  d.offset         = g.f_offset.get();
  d.num_nodes      = g.node_cap;
  d.point_indices  = P.out_pi.get() + g.P0;
  d.point_base     = g.f_base.get();
  d.num_points     = g.n;
  d.xy             = xy;
  d.poly_offsets   = reinterpret_cast<uint32_t const*>(P.poly_offsets + g.PO0);
  d.num_polys      = npg;
  d.ring_offsets   = reinterpret_cast<uint32_t const*>(P.ring_offsets);
  d.num_rings      = static_cast<uint64_t>(P.num_ring_offsets - 1);
  d.ring_edges     = P.ring_edges.get();
  d.lines          = P.lines.get();
  d.ring_amax      = P.ring_amax.get();
  d.poly_x         = P.poly_x;
  d.poly_y         = P.poly_y;
  d.num_vertices   = static_cast<uint64_t>(P.num_vertices);
  d.poly_box       = P.poly_box.get() + g.PO0;
  d.pairs_per_wave = g.pairs_per_wave;
  d.chunk_start    = g.chunk_start.get();
  d.partials       = g.pip_partials.get();
  d.totals         = g.pip_totals.get();
  d.masks          = g.masks.get();
  d.pm_count       = g.pm_count.get();
  d.pm_start       = g.pm_start.get();
  d.pm_cursor      = g.pm_cursor.get();
  d.pm_order       = g.pm_order.get();
  d.pm_units       = g.pm_units.get();
  d.pm_partials    = g.pm_partials.get();
  d.pm_total       = g.pm_total.get();
  d.pm_num_units   = g.pm_num_units.get();
  d.chunk_info     = g.chunk_info.get();
  d.hit_start      = g.hit_start.get();
  d.chunk_cap      = g.cap.chunks;
  d.out_poly       = g.hit_poly.get();
  d.out_pos        = g.hit_pos.get();
  d.hit_cap        = g.cap.hits;
  d.counts         = cnt;
  launch_pip_device(d, s);
  plan_local_hits_kernel<<<grid_for(g.cap.hits, 4096), 256, 0, s>>>(
    g.hit_poly.get(), g.hit_pos.get(), cnt, g.jobs.get(), g.J, P.hit_cap, P.out_poly.get(),
    P.out_point.get());
  QPIP_LAUNCH_CHECK();
  plan_job_first_kernel<<<grid_for(g.J, 256), 256, 0, s>>>(g.hit_pos.get(), cnt, g.jobs.get(),
                                                          g.J, job_first);
  QPIP_LAUNCH_CHECK();
}

// Builds a group's fixed buffers and its polygon boxes (run once, not captured).
void plan_group_setup(BatchPlan& P, PlanGroup& g, hipStream_t s)
{
  std::vector<HostJob> const& hj = P.hj;
  int64_t const* jpo             = P.jpo.data();
  int64_t const* jpoly           = P.jpoly.data();
  g.J   = static_cast<uint32_t>(g.j1 - g.j0);
  g.P0  = jpo[g.j0];
  g.PO0 = jpoly[g.j0];
  g.n   = static_cast<uint64_t>(jpo[g.j1] - g.P0);
  g.npg = static_cast<uint64_t>(jpoly[g.j1] - g.PO0);
  g.key_bits = 1;
  g.maxd     = 0;
  for (int64_t j = g.j0; j < g.j1; ++j) {
    g.key_bits = std::max(g.key_bits, hj[j].key_bits);
    g.maxd     = std::max(g.maxd, hj[j].d);
  }
  g.tree      = g.n > 0;
  g.join      = g.tree && g.npg > 0;
  std::vector<JobDev> hd = job_descs(hj, g.j0, g.j1, jpo, jpoly);
  std::vector<SortTile> ht;
  std::vector<SortSegment> hs;
  group_sort_tiles(g.j0, g.j1, jpo, g.key_bits, ht, hs);
  g.num_stiles = ht.size();
  g.jobs.reset(g.J);
  g.stiles.reset(ht.size());
  g.ssegs.reset(hs.size());
  QPIP_HIP(hipMemcpyAsync(g.jobs.get(), hd.data(), g.J * sizeof(JobDev), hipMemcpyHostToDevice,
                          s));
  if (!ht.empty())
    QPIP_HIP(hipMemcpyAsync(g.stiles.get(), ht.data(), ht.size() * sizeof(SortTile),
                            hipMemcpyHostToDevice, s));
  QPIP_HIP(hipMemcpyAsync(g.ssegs.get(), hs.data(), hs.size() * sizeof(SortSegment),
                          hipMemcpyHostToDevice, s));
  QPIP_HIP(hipStreamSynchronize(s));  // hd, ht, hs are pageable host memory
  if (!g.tree) return;
  g.keys_a.reset(g.n);
  g.keys_b.reset(g.n);
  g.vals_tmp.reset(g.n);
  g.samples.reset((g.n + 31) / 32);
  g.counts0.reset(g.num_stiles * kRadix);
  g.node_cap = forest_capacity(hj, g.j0, g.j1, jpo);
  uint64_t const cap = g.node_cap;
  g.nkey.reset(cap);
  g.nstart.reset(cap);
  g.ncount.reset(cap);
  g.cfirst.reset(cap);
  g.ccount.reset(cap);
  g.njob.reset(cap);
  g.nlevel.reset(cap);
  g.f_internal.reset(cap);
  g.f_length.reset(cap);
  g.f_offset.reset(cap);
  g.f_base.reset(cap);
  g.split.reset(cap);
  g.staging.reset(static_cast<size_t>(g.J) * 64);
  g.trees.reset(g.J);
  g.fmeta.reset(1);
  g.partials.reset(std::max(kLevelBlocks, kScanBlocks) + 1);
  if (!g.join) return;
  uint64_t const npg = g.npg;
  g.boxes.reset(4 * npg);
  double* bminx = g.boxes.get();
  fill_nan_kernel<<<grid_for(4 * npg, 256), 256, 0, s>>>(bminx, 4 * npg);
  QPIP_LAUNCH_CHECK();
  dbuf<uint32_t> prefix(npg + 1);
  NonEmptyOp ne{P.poly_offsets + g.PO0, P.ring_offsets, P.num_ring_offsets, P.num_vertices,
                prefix.get()};
  count_scan_emit<NonEmptyOp, uint32_t, uint32_t>(ne, nullptr, npg, g.partials.get(),
                                                  prefix.get() + npg, s, scan_blocks_for(npg),
                                                  "batch_bbox_scan");
  batch_bbox_kernel<<<grid_for(npg, 256 / kWave), 256, 0, s>>>(
    P.poly_offsets + g.PO0, P.ring_offsets, P.num_ring_offsets, P.poly_x, P.poly_y,
    P.num_vertices, prefix.get(), npg, g.jobs.get(), g.J, bminx, bminx + npg, bminx + 2 * npg,
    bminx + 3 * npg);
  QPIP_LAUNCH_CHECK();
  QPIP_HIP(hipStreamSynchronize(s));  // prefix is freed on return
  g.top_off.reset(g.J);
  g.partials64.reset(kScanBlocks + 1);
  g.jmeta.reset(1);
  g.pip_partials.reset(kScanBlocks + 1);
  g.pip_totals.reset(2);
  g.pm_count.reset(std::max<uint64_t>(g.npg, 1));
  g.pm_start.reset(std::max<uint64_t>(g.npg, 1));
  g.pm_cursor.reset(std::max<uint64_t>(g.npg, 1));
  g.pm_partials.reset(kScanBlocks + 1);
  g.pm_total.reset(1);
  g.pm_num_units.reset(1);
}

// True sizes of a group's last run (valid where no earlier stage overflowed).
PlanCaps plan_measure(PlanGroup const& g, hipStream_t s)
{
  PlanCaps m;
  if (!g.join) return m;
  JoinMeta jm;
  uint64_t tot[2];
  QPIP_HIP(hipMemcpyAsync(&jm, g.jmeta.get(), sizeof(JoinMeta), hipMemcpyDeviceToHost, s));
  QPIP_HIP(hipMemcpyAsync(tot, g.pip_totals.get(), sizeof(tot), hipMemcpyDeviceToHost, s));
  QPIP_HIP(hipStreamSynchronize(s));
  m.frontier = jm.top_pairs > 0 ? 1 : 0;
  for (int l = 0; l < std::max(g.maxd, 1); ++l) {
    m.pairs += jm.level[l].leaf;
    m.frontier = std::max<uint64_t>(m.frontier, jm.level[l].quad);
  }
  m.chunks = tot[0];
  m.hits   = tot[1];
  return m;
}

}  // namespace

}  // namespace qpip

extern "C" int qpip_batched_quadtree_point_in_polygon(int64_t num_jobs,
                                                      const int64_t* job_point_offsets,
                                                      const int64_t* job_polygon_offsets,
                                                      const double* job_bbox,
                                                      const double* job_scale,
                                                      const int8_t* job_max_depth,
                                                      const int32_t* job_max_size,
                                                      const double* points_xy,
                                                      int64_t num_points,
                                                      const int32_t* poly_offsets,
                                                      int64_t num_poly_offsets,
                                                      const int32_t* ring_offsets,
                                                      int64_t num_ring_offsets,
                                                      const double* poly_x,
                                                      const double* poly_y,
                                                      int64_t num_poly_x,
                                                      int64_t num_poly_y,
                                                      const qpip_resource* mr,
                                                      void* stream_,
                                                      uint32_t** point_indices,
                                                      uint32_t** polygon_index,
                                                      uint32_t** point_index,
                                                      int64_t* job_hit_offsets,
                                                      int64_t* num_hits)
{
  using namespace qpip;
  try {
    QPIP_EXPECTS(point_indices && polygon_index && point_index && job_hit_offsets && num_hits,
                 "output pointers must not be null");
    *point_indices = *polygon_index = *point_index = nullptr;
    *num_hits                                      = 0;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    QPIP_EXPECTS(num_jobs >= 0, "number of jobs must not be negative");
    if (num_jobs == 0) {
      job_hit_offsets[0] = 0;
      return QPIP_SUCCESS;
    }
    std::vector<HostJob> hj =
      prepare_jobs(num_jobs, job_point_offsets, job_polygon_offsets, job_bbox, job_scale,
                   job_max_depth, job_max_size, num_points, num_poly_offsets, num_poly_x,
                   num_poly_y);
    if (num_points > 0) QPIP_EXPECTS(points_xy, "points must not be null");

    StreamScope scope(stream);
    OutputAllocator oa(mr, stream);
    uint32_t* o_pi = num_points > 0 ? oa.make<uint32_t>(num_points) : nullptr;
    std::vector<uint64_t> job_hits(num_jobs, 0);
    std::vector<GroupResult> groups;
    for (auto const& [j0, j1] : job_groups(hj, job_point_offsets)) {
      groups.emplace_back();
      run_group(hj, j0, j1, job_point_offsets, job_polygon_offsets, points_xy, poly_offsets,
                ring_offsets, num_ring_offsets, poly_x, poly_y, num_poly_x,
                o_pi ? o_pi + job_point_offsets[j0] : nullptr, groups.back(), job_hits, stream);
    }
    uint64_t H = 0;
    job_hit_offsets[0] = 0;
    for (int64_t j = 0; j < num_jobs; ++j) {
      H += job_hits[j];
      job_hit_offsets[j + 1] = static_cast<int64_t>(H);
    }
    uint32_t* o_poly  = H > 0 ? oa.make<uint32_t>(H) : nullptr;
    uint32_t* o_point = H > 0 ? oa.make<uint32_t>(H) : nullptr;
    uint64_t at = 0;
    for (auto& g : groups) {
      if (g.hits == 0) continue;
      QPIP_HIP(hipMemcpyAsync(o_poly + at, g.poly.get(), g.hits * 4, hipMemcpyDeviceToDevice, stream));
      QPIP_HIP(hipMemcpyAsync(o_point + at, g.point.get(), g.hits * 4, hipMemcpyDeviceToDevice, stream));
      at += g.hits;
    }
    oa.keep();
    *point_indices = o_pi;
    *polygon_index = o_poly;
    *point_index   = o_point;
    *num_hits      = static_cast<int64_t>(H);
    return QPIP_SUCCESS;
  } catch (...) {
    return qpip::error_status();
  }
}

// ---------------------------------------------------------------------------------
// Captured plan entry points
// ---------------------------------------------------------------------------------
namespace {

qpip::BatchPlan* plan_of(qpip_batch_plan* p) { return reinterpret_cast<qpip::BatchPlan*>(p); }

void check_device(qpip::BatchPlan const& P)
{
  int dev = 0;
  QPIP_HIP(hipGetDevice(&dev));
  QPIP_EXPECTS(dev == P.device, "the plan belongs to another device");
}

}  // namespace

extern "C" int qpip_batch_plan_create(int64_t num_jobs,
                                      const int64_t* job_point_offsets,
                                      const int64_t* job_polygon_offsets,
                                      const double* job_bbox,
                                      const double* job_scale,
                                      const int8_t* job_max_depth,
                                      const int32_t* job_max_size,
                                      const double* points_xy,
                                      int64_t num_points,
                                      const int32_t* poly_offsets,
                                      int64_t num_poly_offsets,
                                      const int32_t* ring_offsets,
                                      int64_t num_ring_offsets,
                                      const double* poly_x,
                                      const double* poly_y,
                                      int64_t num_poly_x,
                                      int64_t num_poly_y,
                                      double headroom,
                                      void* stream_,
                                      qpip_batch_plan** plan)
{
  using namespace qpip;
  return qpip::guarded([&] {
    QPIP_EXPECTS(plan != nullptr, "output pointers must not be null");
    *plan = nullptr;
    QPIP_EXPECTS(num_jobs >= 1, "a plan needs at least one job");
    QPIP_EXPECTS(headroom >= 1.0 && headroom <= 64.0, "headroom must be in [1, 64]");
    std::vector<HostJob> hj =
      prepare_jobs(num_jobs, job_point_offsets, job_polygon_offsets, job_bbox, job_scale,
                   job_max_depth, job_max_size, num_points, num_poly_offsets, num_poly_x,
                   num_poly_y);
    if (num_points > 0) QPIP_EXPECTS(points_xy, "points must not be null");
    QPIP_EXPECTS(num_ring_offsets >= 1 || num_poly_offsets <= 1,
                 "ring offsets need at least one entry");
    // the caller's pending writes of the points and polygons come first
    QPIP_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream_)));

    auto P = std::make_unique<BatchPlan>();
    QPIP_HIP(hipGetDevice(&P->device));
    QPIP_HIP(hipStreamCreateWithFlags(&P->cstream, hipStreamNonBlocking));
    hipStream_t const s = P->cstream;
    StreamScope scope(s);
    P->hj               = hj;
    P->jpo.assign(job_point_offsets, job_point_offsets + num_jobs + 1);
    P->jpoly.assign(job_polygon_offsets, job_polygon_offsets + num_jobs + 1);
    P->num_jobs         = num_jobs;
    P->num_points       = num_points;
    P->points_xy        = points_xy;
    P->poly_offsets     = poly_offsets;
    P->ring_offsets     = ring_offsets;
    P->num_ring_offsets = num_ring_offsets;
    P->num_vertices     = num_poly_x;
    P->poly_x           = poly_x;
    P->poly_y           = poly_y;
    int64_t const rings = num_ring_offsets > 0 ? num_ring_offsets - 1 : 0;
    P->ring_edges.reset(std::max<int64_t>(rings, 1));
    P->lines.reset(std::max<int64_t>(num_poly_x, 1));
    P->ring_amax.reset(std::max<int64_t>(rings, 1));
    uint64_t const polys = num_poly_offsets > 1 ? static_cast<uint64_t>(num_poly_offsets - 1) : 0;
    P->poly_box.reset(std::max<uint64_t>(polys, 1));
    {
      dbuf<PolyBox> ring_box(std::max<int64_t>(rings, 1));
      if (rings > 0) {
        build_edges_kernel<<<grid_for(static_cast<uint64_t>(rings), 256 / kWave), 256, 0, s>>>(
          reinterpret_cast<uint32_t const*>(ring_offsets), rings, poly_x, poly_y, num_poly_x,
          nullptr, nullptr, P->ring_edges.get(), P->lines.get(),
          P->ring_amax.get(), ring_box.get());
        QPIP_LAUNCH_CHECK();
      }
      if (polys > 0) {
        poly_box_kernel<<<grid_for(polys, 256), 256, 0, s>>>(
          reinterpret_cast<uint32_t const*>(poly_offsets), polys, static_cast<uint64_t>(rings),
          ring_box.get(), P->poly_box.get());
        QPIP_LAUNCH_CHECK();
      }
      QPIP_HIP(hipStreamSynchronize(s));  // ring_box goes back to the pool
    }
    for (auto const& [j0, j1] : job_groups(P->hj, P->jpo.data())) {
      auto g = std::make_unique<PlanGroup>();
      g->j0  = j0;
      g->j1  = j1;
      plan_group_setup(*P, *g, s);
      if (g->join) {
        uint64_t const c = std::max<uint64_t>(uint64_t{1} << 16, 8 * g->npg);
        g->cap           = PlanCaps{c, c, 2 * c, std::max<uint64_t>(uint64_t{1} << 16, g->n)};
      }
      P->groups.push_back(std::move(g));
    }
    size_t const G = P->groups.size();
    P->counts.reset(G);
    P->job_first.reset(static_cast<size_t>(num_jobs));
    P->out_pi.reset(static_cast<size_t>(num_points));

    // sizing runs: grow each exceeded capacity until one run fits
    auto alloc_all = [&] {
      uint64_t H = 0;
      for (auto& g : P->groups) {
        if (!g->join) continue;
        plan_alloc(*g);
        g->pairs_per_wave = pip_pairs_per_wave(g->cap.pairs);
        H += g->cap.hits;
      }
      P->hit_cap = H;
      P->out_poly.reset(std::max<uint64_t>(H, 1));
      P->out_point.reset(std::max<uint64_t>(H, 1));
    };
    std::vector<DeviceCounts> hc(G);
    std::vector<PlanCaps> measured(G);
    bool fits = false;
    for (int attempt = 0; attempt < 16 && !fits; ++attempt) {
      alloc_all();
      for (size_t gi = 0; gi < G; ++gi) plan_enqueue(*P, *P->groups[gi], gi, s);
      QPIP_HIP(hipMemcpyAsync(hc.data(), P->counts.get(), G * sizeof(DeviceCounts),
                              hipMemcpyDeviceToHost, s));
      QPIP_HIP(hipStreamSynchronize(s));
      fits = true;
      for (size_t gi = 0; gi < G; ++gi) {
        if (hc[gi].flags & 1u)
          throw logic_error("cuSpatial failure: forest child range or index out of range");
        PlanGroup& g = *P->groups[gi];
        measured[gi] = plan_measure(g, s);
        if (!(hc[gi].flags & 2u)) continue;
        fits          = false;
        PlanCaps& c   = g.cap;
        PlanCaps const& m = measured[gi];
        bool grown    = false;
        auto grow     = [&](uint64_t& cap, uint64_t need) {
          if (need <= cap) return;
          cap   = need + need / 4 + 1024;
          grown = true;
        };
        grow(c.pairs, m.pairs);
        grow(c.frontier, m.frontier);
        grow(c.chunks, m.chunks);
        grow(c.hits, m.hits);
        if (!grown) {  // a later stage saw clamped input: widen everything
          c.pairs *= 2;
          c.frontier *= 2;
          c.chunks *= 2;
          c.hits *= 2;
        }
      }
    }
    if (!fits) throw device_error("batch plan sizing did not converge");
    // final capacities: the sizing run's counts times the headroom
    for (size_t gi = 0; gi < G; ++gi) {
      PlanGroup& g = *P->groups[gi];
      if (!g.join) continue;
      PlanCaps const& m = measured[gi];
      auto room         = [&](uint64_t v) {
        return static_cast<uint64_t>(static_cast<double>(v) * headroom) + 1024;
      };
      g.cap = PlanCaps{room(m.pairs), room(m.frontier), room(m.chunks), room(m.hits)};
      QPIP_EXPECTS(g.cap.pairs < (uint64_t{1} << 32) && g.cap.chunks < (uint64_t{1} << 31) &&
                     g.cap.hits < (uint64_t{1} << 32),
                   "batch plan capacities exceed 32-bit row counts");
    }
    alloc_all();
    for (size_t gi = 0; gi < G; ++gi)
      if (P->groups[gi]->join) P->groups[gi]->pairs_per_wave = pip_pairs_per_wave(measured[gi].pairs);
    QPIP_HIP(hipStreamSynchronize(s));

    // capture the whole sequence once
    QPIP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    set_capturing(true);
    try {
      for (size_t gi = 0; gi < G; ++gi) plan_enqueue(*P, *P->groups[gi], gi, s);
    } catch (...) {
      set_capturing(false);
      hipGraph_t broken = nullptr;
      (void)hipStreamEndCapture(s, &broken);
      if (broken) (void)hipGraphDestroy(broken);
      throw;
    }
    set_capturing(false);
    QPIP_HIP(hipStreamEndCapture(s, &P->graph));
    QPIP_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    *plan = reinterpret_cast<qpip_batch_plan*>(P.release());
  });
}

extern "C" int qpip_batch_plan_launch(qpip_batch_plan* plan, void* stream_)
{
  using namespace qpip;
  return qpip::guarded([&] {
    QPIP_EXPECTS(plan != nullptr, "plan must not be null");
    BatchPlan& P = *plan_of(plan);
    check_device(P);
    hipStream_t const s = static_cast<hipStream_t>(stream_);
    if (P.launched && P.last_stream != s) QPIP_HIP(hipStreamSynchronize(P.last_stream));
    QPIP_HIP(hipGraphLaunch(P.exec, s));
    P.last_stream = s;
    P.launched    = true;
  });
}

extern "C" int qpip_batch_plan_result(qpip_batch_plan* plan,
                                      const qpip_resource* mr,
                                      void* stream_,
                                      uint32_t** point_indices,
                                      uint32_t** polygon_index,
                                      uint32_t** point_index,
                                      int64_t* job_hit_offsets,
                                      int64_t* num_hits)
{
  using namespace qpip;
  return qpip::guarded([&] {
    QPIP_EXPECTS(plan != nullptr, "plan must not be null");
    QPIP_EXPECTS(point_indices && polygon_index && point_index && job_hit_offsets && num_hits,
                 "output pointers must not be null");
    BatchPlan& P = *plan_of(plan);
    check_device(P);
    QPIP_EXPECTS(P.launched, "the plan has not been launched");
    hipStream_t const rs = static_cast<hipStream_t>(stream_);
    QPIP_HIP(hipStreamSynchronize(rs));
    QPIP_HIP(hipStreamSynchronize(P.last_stream));
    size_t const G = P.groups.size();
    std::vector<DeviceCounts> hc(G);
    std::vector<int64_t> first(static_cast<size_t>(P.num_jobs));
    // Copies on the caller's stream, never the legacy stream: another thread may be
    // capturing a plan, and a legacy-stream copy would have to wait on its stream.
    QPIP_HIP(hipMemcpyAsync(hc.data(), P.counts.get(), G * sizeof(DeviceCounts),
                            hipMemcpyDeviceToHost, rs));
    QPIP_HIP(hipMemcpyAsync(first.data(), P.job_first.get(), first.size() * sizeof(int64_t),
                            hipMemcpyDeviceToHost, rs));
    QPIP_HIP(hipStreamSynchronize(rs));
    uint32_t flags = 0;
    for (auto const& c : hc) flags |= c.flags;
    if (flags & 1u) throw logic_error("cuSpatial failure: forest child range or index out of range");
    if (flags & 2u)
      throw capacity_error(
        "batch plan capacity exceeded: this launch produced more pairs, chunks or hits than "
        "the plan was sized for; create the plan again (or with more headroom)");
    uint64_t const H = G ? hc[G - 1].hit_base + hc[G - 1].hits : 0;
    for (int64_t j = 0; j < P.num_jobs; ++j) job_hit_offsets[j] = first[j];
    job_hit_offsets[P.num_jobs] = static_cast<int64_t>(H);
    *num_hits                   = static_cast<int64_t>(H);
    if (!mr) {  // the plan's own buffers
      *point_indices = P.out_pi.get();
      *polygon_index = P.out_poly.get();
      *point_index   = P.out_point.get();
      return;
    }
    hipStream_t const s = static_cast<hipStream_t>(stream_);
    StreamScope scope(s);
    OutputAllocator oa(mr, s);
    uint64_t const n = static_cast<uint64_t>(P.num_points);
    uint32_t* o_pi   = oa.make<uint32_t>(n);
    uint32_t* o_poly = oa.make<uint32_t>(H);
    uint32_t* o_pt   = oa.make<uint32_t>(H);
    if (n) QPIP_HIP(hipMemcpyAsync(o_pi, P.out_pi.get(), n * 4, hipMemcpyDeviceToDevice, s));
    if (H) {
      QPIP_HIP(hipMemcpyAsync(o_poly, P.out_poly.get(), H * 4, hipMemcpyDeviceToDevice, s));
      QPIP_HIP(hipMemcpyAsync(o_pt, P.out_point.get(), H * 4, hipMemcpyDeviceToDevice, s));
    }
    oa.keep();
    *point_indices = o_pi;
    *polygon_index = o_poly;
    *point_index   = o_pt;
  });
}

extern "C" int qpip_batch_plan_destroy(qpip_batch_plan* plan)
{
  using namespace qpip;
  return qpip::guarded([&] { delete plan_of(plan); });
}
