// The point-in-polygon stage as a fixed launch sequence whose sizes live on the
// device (for captured launch sequences: no host synchronisation, no allocation).
#pragma once

#include "qpip_edges.hpp"

namespace qpip {

// Per launch, on the device. `pairs` is written by the producer of the pairs
// before the stage runs; `chunks` and `hits` by the stage (each clamped to its
// capacity); `flags` bit 1: an index out of range, bit 2: a capacity exceeded.
struct DeviceCounts {
  uint64_t pairs;
  uint64_t chunks;
  uint64_t hits;
  uint64_t hit_base;  // first output row of this part of a larger result
  uint32_t flags;
  uint32_t pad;
};

struct PipDevice {
  // pairs (polygon, leaf) in leaf-offset order, at most pair_cap of them
  uint32_t const* pair_poly;
  uint32_t const* pair_quad;
  uint64_t pair_cap;
  // the quadtree (num_nodes entries may be a capacity)
  uint32_t const* length;
  uint32_t const* offset;
  uint64_t num_nodes;
  uint32_t const* point_indices;
  uint64_t num_points;
  double const* xy;  // interleaved points
  // non-null: point_indices are local to each node's job; point_base[node] is
  // the job's first point
  uint32_t const* point_base;
  // polygons and their edge table
  uint32_t const* poly_offsets;
  uint64_t num_polys;
  uint32_t const* ring_offsets;
  uint64_t num_rings;
  uint32_t const* ring_edges;
  LineEdge const* lines;    // per edge: the test's line form
  double const* ring_amax;  // per ring: largest |vertex coordinate|
  double const* poly_x  = nullptr;  // the vertices (the exact retest walks them)
  double const* poly_y  = nullptr;
  uint64_t num_vertices = 0;
  PolyBox const* poly_box;  // per polygon (null: the test filters no points)
  uint32_t pairs_per_wave;
  // workspace
  uint32_t* chunk_start;  // pair_cap + 1
  uint64_t* partials;     // kScanBlocks + 1
  uint64_t* totals;       // 2: chunk and hit totals before clamping
  uint64_t* masks;        // chunk_cap
  uint2* chunk_info;      // chunk_cap
  uint64_t* hit_start;    // chunk_cap
  uint64_t chunk_cap;
  // the polygon-major test's order (null pm_order: the leaf-major test)
  uint32_t* pm_count;     // num_polys
  uint32_t* pm_start;     // num_polys
  uint32_t* pm_cursor;    // num_polys
  uint4* pm_order;        // pair_cap: per order position, the pair's record
  uint4* pm_units;        // pair_cap
  uint64_t* pm_partials;  // kScanBlocks + 1
  uint64_t* pm_total;     // 1
  uint64_t* pm_num_units; // 1
  // output: (polygon, point position) hits, at most hit_cap
  uint32_t* out_poly;
  uint32_t* out_pos;
  uint64_t hit_cap;
  DeviceCounts* counts;
};

// Buffers of the polygon-major test's pair order.
struct PmBuffers {
  uint32_t* count;      // per polygon
  uint32_t* start;      // per polygon
  uint32_t* cursor;     // per polygon
  uint4* order;         // per pair: its record (leaf count, first, base, slot)
  uint4* units;         // per pair (at most one unit per pair)
  uint64_t* partials;   // kScanBlocks + 1
  uint64_t* total;      // 1
  uint64_t* num_units;  // 1
};

// Enqueues the stage on `stream`: chunks, test, scan, write.
void launch_pip_device(PipDevice const& d, hipStream_t stream);

// The quadtree PIP join over host-visible sizes (qpip_quadtree_point_in_polygon and
// its _xy form; x/y or xy). With point_base non-null, point_indices are local to
// each node's job and point_base[node] is that job's first point (the batched
// eager call); the C entry points pass null.
int quadtree_point_in_polygon_entry(const uint32_t* poly_index, const uint32_t* quad_index,
                                    int64_t num_pairs, const uint32_t* key, const uint8_t* level,
                                    const uint8_t* is_internal_node, const uint32_t* length,
                                    const uint32_t* offset, int64_t num_nodes,
                                    const uint32_t* point_indices, int64_t num_point_indices,
                                    const double* x, const double* y, const double* xy,
                                    int64_t num_x, int64_t num_y, const uint32_t* poly_offsets,
                                    int64_t num_poly_offsets, const uint32_t* ring_offsets,
                                    int64_t num_ring_offsets, const double* poly_x,
                                    const double* poly_y, int64_t num_poly_x, int64_t num_poly_y,
                                    const qpip_resource* mr, void* stream,
                                    uint32_t** polygon_index, uint32_t** point_index,
                                    int64_t* num_hits, const uint32_t* point_base);

// The test kernel's pairs-per-wave segment length for a join of num_pairs pairs
// (the runtime tuning parameters pip_min_pairs_per_wave, pip_pairs_shift).
uint32_t pip_pairs_per_wave(uint64_t num_pairs);

}  // namespace qpip
