/*
 * cuspatial_amd — C ABI of the MI355X-native quadtree point-in-polygon join.
 *
 * Each entry point replaces one function of the reference's column API
 * (rapidsai/cuspatial 25.06, cited as file:line relative to the reference root):
 * the same inputs, in the same order and meaning, as plain device pointers and
 * element counts instead of cudf::column_view, and the same output columns as
 * device buffers instead of cudf::column.
 *
 * Conventions
 *  - All array arguments are device pointers (HBM). Coordinates are float64;
 *    the reference's float32 instantiation is not provided.
 *  - `stream` is a hipStream_t (NULL: the default stream). Entry points are
 *    stream-ordered, as the reference's header-only API is: the work is queued
 *    on `stream`, and the outputs are ready for any work queued on `stream`
 *    after the call returns. The host waits only where a data-dependent size
 *    must be known to allocate an output (at most once per call of the
 *    quadtree, bounding-box, join and point-in-polygon entry points; a second
 *    time when a join's candidates outgrow the first guess of their count).
 *    Synchronise `stream` before reading outputs from the host.
 *  - Outputs are allocated from `mr` when it is non-NULL (the counterpart of the
 *    reference's `rmm::device_async_resource_ref mr` argument); with mr == NULL
 *    they come from the library's pool and are released with qpip_free().
 *    Zero-length outputs are returned as NULL with a count of 0.
 *  - Return value: QPIP_SUCCESS, or an error code with a message available from
 *    qpip_last_error_message() on the calling thread. QPIP_ERROR_LOGIC
 *    corresponds to cuspatial::logic_error (same reason strings),
 *    QPIP_ERROR_DEVICE to cuspatial::cuda_error, QPIP_ERROR_OOM to
 *    rmm::out_of_memory.
 */
#ifndef CUSPATIAL_AMD_H
#define CUSPATIAL_AMD_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define QPIP_API __attribute__((visibility("default")))
#else
#define QPIP_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define QPIP_SUCCESS      0
#define QPIP_ERROR_LOGIC  1
#define QPIP_ERROR_DEVICE 2
#define QPIP_ERROR_OOM    3
#define QPIP_ERROR_CAPACITY 4 /* a launch outgrew a capacity fixed when a plan was made */

/* Caller-supplied device memory resource for outputs. */
typedef struct qpip_resource {
  void* (*allocate)(void* ctx, size_t bytes, void* stream);
  void (*deallocate)(void* ctx, void* ptr, size_t bytes, void* stream);
  void* ctx;
} qpip_resource;

/* Message of the last failed call on this thread ("" if none). */
QPIP_API const char* qpip_last_error_message(void);

/* Library version string. */
QPIP_API const char* qpip_version(void);

/* Releases a buffer returned with mr == NULL. NULL is ignored. The library's
 * own writes to it are ordered before any reuse; the caller's uses must have
 * finished (as with hipFree). */
QPIP_API void qpip_free(void* ptr);

/* Caps the bytes the library's pool may hold in live buffers (0: no cap).
 * Allocations beyond the cap fail with QPIP_ERROR_OOM. */
QPIP_API void qpip_set_memory_limit(size_t bytes);

/* Returns cached, unused pool memory to the device. */
QPIP_API void qpip_trim(void);

/* Bytes the library's pool holds on the current device: in live buffers
 * (temporaries of running calls and pool-allocated outputs) and cached for
 * reuse. No counterpart in the reference (RMM reports its own pools); a server
 * admitting work by free device memory counts the cached bytes as free. */
QPIP_API void qpip_memory_info(size_t* live_bytes, size_t* cached_bytes);

/* Tuning parameters (no counterpart in the reference; set by the autotuner,
 * cuspatial_amd/autotune.py). Process-wide, read when a kernel is launched:
 *   pip_min_pairs_per_wave  shortest segment of a leaf's pairs that one wave of
 *                           the point-in-polygon test takes (default 2)
 *   pip_pairs_shift         segments are at most max(min, pairs >> shift) pairs
 *                           long (default 24)
 *   pip_waves_per_block     waves per workgroup of that test: 2, 4 or 8 (values
 *                           round up to one of these; default 2)
 *   pip_stage_edges         ring edges a wave stages in LDS per batch: 64 or 128
 *                           (values above 64 select 128; default 64)
 *   pip_xcd_run             consecutive workgroups of that test each XCD takes
 *                           in turn (1: the dispatch order; default 16)
 *   pip_box_filter          1: that test skips points outside a polygon's vertex
 *                           box and packs the rest into fewer lanes; 0: it tests
 *                           every candidate point (default 1)
 *   pip_polygon_group       0: the test walks the pairs leaf by leaf (default);
 *                           n > 0: it groups them by polygon, one wave taking up
 *                           to n pairs of one polygon and packing their kept
 *                           points from several leaves into full waves
 * Results do not depend on them, only the time. Errors: unknown name, value out
 * of range. */
QPIP_API int qpip_tuning_set(const char* name, int64_t value);
QPIP_API int qpip_tuning_get(const char* name, int64_t* value);

/* Kernel timing (no counterpart in the reference; for benchmarks). While
 * enabled, every kernel launch is bracketed by HIP events on its stream.
 * qpip_profile_report writes {"kernel name": [launches, total ms], ...} as JSON
 * into buf (truncated to len) and returns the full length. */
QPIP_API void qpip_profile_enable(int on);
QPIP_API void qpip_profile_reset(void);
QPIP_API int64_t qpip_profile_report(char* buf, size_t len);

/*
 * quadtree_on_points
 * Replaces cuspatial::quadtree_on_points
 *   (cpp/include/cuspatial/point_quadtree.hpp:68-78, cpp/src/indexing/point_quadtree.cu:154-180;
 *    algorithm detail/point_quadtree.cuh:238-272).
 * Inputs x, y (num_x, num_y points). Clamps: bbox ordered, max_size >= 1,
 * 0 <= max_depth <= 15, scale >= max(dx, dy) / ((1 << max_depth) + 2).
 * Outputs: point_indices (UINT32, num_x), and the five quadtree columns
 * key (UINT32), level (UINT8), is_internal_node (BOOL8 as uint8), length (UINT32),
 * offset (UINT32), each *num_nodes long.
 * Errors: "x and y columns must have the same length".
 */
QPIP_API int qpip_quadtree_on_points(const double* x,
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
                            int64_t* num_nodes);

/*
 * quadtree_on_points over interleaved points
 * Replaces the header-only cuspatial::quadtree_on_points(PointIterator points_first,
 *   PointIterator points_last, vec_2d<T> vertex_1, vec_2d<T> vertex_2, T scale,
 *   int8_t max_depth, int32_t max_size) (cpp/include/cuspatial/point_quadtree.cuh:170-180),
 *   whose points are vec_2d<double>: interleaved x, y pairs, the layout in which a
 *   point GeoSeries holds its coordinates (python/cuspatial/cuspatial/core/geoseries.py:237-250).
 * points_xy holds 2 * num_points doubles; everything else is as for
 * qpip_quadtree_on_points.
 */
QPIP_API int qpip_quadtree_on_points_xy(const double* points_xy,
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
                                        int64_t* num_nodes);

/*
 * polygon_bounding_boxes
 * Replaces cuspatial::polygon_bounding_boxes
 *   (cpp/include/cuspatial/bounding_boxes.hpp:79-85,
 *    cpp/src/bounding_boxes/polygon_bounding_boxes.cu:132-159;
 *    algorithm detail/bounding_boxes.cuh:62-90,139-184).
 * poly_offsets / ring_offsets are GeoArrow INT32 offsets (num_polygons + 1 and
 * num_rings + 1 entries). Outputs minx, miny, maxx, maxy, *num_boxes =
 * num_poly_offsets - 1 rows (0 when num_poly_offsets <= 1).
 * Errors: "x and y must be the same size".
 */
QPIP_API int qpip_polygon_bounding_boxes(const int32_t* poly_offsets,
                                int64_t num_poly_offsets,
                                const int32_t* ring_offsets,
                                int64_t num_ring_offsets,
                                const double* x,
                                const double* y,
                                int64_t num_x,
                                int64_t num_y,
                                double expansion_radius,
                                const qpip_resource* mr,
                                void* stream,
                                double** minx,
                                double** miny,
                                double** maxx,
                                double** maxy,
                                int64_t* num_boxes);

/*
 * join_quadtree_and_bounding_boxes
 * Replaces cuspatial::join_quadtree_and_bounding_boxes
 *   (cpp/include/cuspatial/spatial_join.hpp:66-75, cpp/src/join/quadtree_bbox_filtering.cu:90-126;
 *    algorithm detail/join/quadtree_bbox_filtering.cuh:37-188).
 * The quadtree is the five columns returned by qpip_quadtree_on_points; the
 * bounding boxes are four columns of num_bboxes rows.
 * Outputs bbox_offset, quad_offset (UINT32, *num_pairs), ordered by the leaf's
 * first point position, then by traversal order.
 * Errors: "scale must be positive", "invalid bounding box (x_min, x_max, y_min, y_max)",
 * "maximum depth must be positive and less than 16".
 */
QPIP_API int qpip_join_quadtree_and_bounding_boxes(const uint32_t* key,
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
                                          void* stream,
                                          uint32_t** bbox_offset,
                                          uint32_t** quad_offset,
                                          int64_t* num_pairs);

/*
 * quadtree_point_in_polygon
 * Replaces cuspatial::quadtree_point_in_polygon
 *   (cpp/include/cuspatial/spatial_join.hpp:116-126, cpp/src/join/quadtree_point_in_polygon.cu:143-191;
 *    algorithm detail/join/quadtree_point_in_polygon.cuh:104-235).
 * poly_quad_pairs = (poly_index, quad_index) from the join; the quadtree's
 * length/offset columns (the other three are not read and may be NULL);
 * point_indices from qpip_quadtree_on_points; points x, y; polygons as
 * GeoArrow offsets read as uint32 (as the reference does) and vertices.
 * Outputs polygon_index, point_index (UINT32, *num_hits); point_index is a
 * position in point_indices. Order: pair order, then point order in the leaf.
 * Errors: "number of points must be the same for both x and y columns",
 * "numbers of vertices must be the same for both x and y columns".
 */
QPIP_API int qpip_quadtree_point_in_polygon(const uint32_t* poly_index,
                                   const uint32_t* quad_index,
                                   int64_t num_pairs,
                                   const uint32_t* key,
                                   const uint8_t* level,
                                   const uint8_t* is_internal_node,
                                   const uint32_t* length,
                                   const uint32_t* offset,
                                   int64_t num_nodes,
                                   const uint32_t* point_indices,
                                   int64_t num_point_indices,
                                   const double* x,
                                   const double* y,
                                   int64_t num_x,
                                   int64_t num_y,
                                   const uint32_t* poly_offsets,
                                   int64_t num_poly_offsets,
                                   const uint32_t* ring_offsets,
                                   int64_t num_ring_offsets,
                                   const double* poly_x,
                                   const double* poly_y,
                                   int64_t num_poly_x,
                                   int64_t num_poly_y,
                                   const qpip_resource* mr,
                                   void* stream,
                                   uint32_t** polygon_index,
                                   uint32_t** point_index,
                                   int64_t* num_hits);

/*
 * quadtree_point_in_polygon over interleaved points
 * Replaces the header-only cuspatial::quadtree_point_in_polygon(poly_indices_first,
 *   poly_indices_last, quad_indices_first, quadtree, point_indices_first,
 *   point_indices_last, PointIterator points_first, polygons)
 *   (cpp/include/cuspatial/spatial_join.cuh:116-130), whose points are vec_2d<double>
 *   (interleaved x, y pairs). points_xy holds 2 * num_points doubles; everything
 *   else is as for qpip_quadtree_point_in_polygon.
 * Errors: "number of points must be the same for both x and y columns" when
 * num_point_indices != num_points.
 */
QPIP_API int qpip_quadtree_point_in_polygon_xy(const uint32_t* poly_index,
                                               const uint32_t* quad_index,
                                               int64_t num_pairs,
                                               const uint32_t* key,
                                               const uint8_t* level,
                                               const uint8_t* is_internal_node,
                                               const uint32_t* length,
                                               const uint32_t* offset,
                                               int64_t num_nodes,
                                               const uint32_t* point_indices,
                                               int64_t num_point_indices,
                                               const double* points_xy,
                                               int64_t num_points,
                                               const uint32_t* poly_offsets,
                                               int64_t num_poly_offsets,
                                               const uint32_t* ring_offsets,
                                               int64_t num_ring_offsets,
                                               const double* poly_x,
                                               const double* poly_y,
                                               int64_t num_poly_x,
                                               int64_t num_poly_y,
                                               const qpip_resource* mr,
                                               void* stream,
                                               uint32_t** polygon_index,
                                               uint32_t** point_index,
                                               int64_t* num_hits);

/*
 * point_in_polygon (brute force, every point against every polygon)
 * Replaces cuspatial::point_in_polygon
 *   (cpp/include/cuspatial/point_in_polygon.hpp:75-82,
 *    cpp/src/point_in_polygon/point_in_polygon.cu:105-148;
 *    algorithm detail/point_in_polygon.cuh:42-102).
 * Outputs one INT32 per point (*num_results = num_x): bit j is set when the point
 * is inside polygon j (same test as the quadtree path). At most 31 polygons.
 * Errors: "All points must have both x and y values",
 * "Number of polygons cannot exceed 31".
 */
QPIP_API int qpip_point_in_polygon(const double* x,
                                   const double* y,
                                   int64_t num_x,
                                   int64_t num_y,
                                   const int32_t* poly_offsets,
                                   int64_t num_poly_offsets,
                                   const int32_t* ring_offsets,
                                   int64_t num_ring_offsets,
                                   const double* poly_x,
                                   const double* poly_y,
                                   int64_t num_poly_x,
                                   int64_t num_poly_y,
                                   const qpip_resource* mr,
                                   void* stream,
                                   int32_t** result,
                                   int64_t* num_results);

/*
 * point_in_polygon without allocation. No counterpart in the reference: a variant of
 * qpip_point_in_polygon for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that must not call hipMalloc, harnesses that
 * forbid it). Same checks, test and output as qpip_point_in_polygon, but the caller
 * passes the output, `result` (num_x INT32 in device memory), and the edge table
 * lives in a fixed device array of the library (8.75 MiB per device, kept until
 * the process ends). The call allocates nothing, queues its two kernels on `stream`
 * and returns without waiting.
 * Limits: at most QPIP_NOALLOC_MAX_VERTICES polygon vertices and
 * QPIP_NOALLOC_MAX_RINGS rings; larger inputs fail with QPIP_ERROR_CAPACITY.
 * Every call on one GPU uses that GPU's one scratch array: calls that overlap on
 * a GPU (other streams, other threads) race and must be serialised by the caller.
 */
#define QPIP_NOALLOC_MAX_VERTICES 262144
#define QPIP_NOALLOC_MAX_RINGS    65536
QPIP_API int qpip_point_in_polygon_noalloc(const double* x,
                                           const double* y,
                                           int64_t num_x,
                                           int64_t num_y,
                                           const int32_t* poly_offsets,
                                           int64_t num_poly_offsets,
                                           const int32_t* ring_offsets,
                                           int64_t num_ring_offsets,
                                           const double* poly_x,
                                           const double* poly_y,
                                           int64_t num_poly_x,
                                           int64_t num_poly_y,
                                           void* stream,
                                           int32_t* result);

/*
 * pairwise_point_in_polygon (point i against polygon i)
 * Replaces cuspatial::pairwise_point_in_polygon
 *   (cpp/include/cuspatial/point_in_polygon.hpp:124-131,
 *    cpp/src/point_in_polygon/point_in_polygon.cu:105-148;
 *    algorithm detail/point_in_polygon.cuh:104-160).
 * Outputs one UINT8 (0/1) per pair.
 * Errors: "All points must have both x and y values",
 * "Must pass in the same number of points as polygons.".
 */
QPIP_API int qpip_pairwise_point_in_polygon(const double* x,
                                            const double* y,
                                            int64_t num_x,
                                            int64_t num_y,
                                            const int32_t* poly_offsets,
                                            int64_t num_poly_offsets,
                                            const int32_t* ring_offsets,
                                            int64_t num_ring_offsets,
                                            const double* poly_x,
                                            const double* poly_y,
                                            int64_t num_poly_x,
                                            int64_t num_poly_y,
                                            const qpip_resource* mr,
                                            void* stream,
                                            uint8_t** result,
                                            int64_t* num_results);

/*
 * Batched quadtree point-in-polygon (no counterpart in the reference's API: it
 * runs the reference's four-call sequence for many independent jobs at once).
 * For every job j the results are exactly those of
 *   qpip_quadtree_on_points_xy -> qpip_polygon_bounding_boxes ->
 *   qpip_join_quadtree_and_bounding_boxes -> qpip_quadtree_point_in_polygon_xy
 * on the job's points and polygons with the job's bbox, scale, max_depth and
 * max_size, after the Python layer's bbox ordering and scale clamp
 * (python/cuspatial/cuspatial/core/spatial/indexing.py:165-189, join.py:144-156).
 * Job descriptions are HOST arrays: job_point_offsets and job_polygon_offsets
 * (num_jobs + 1 each; job j owns points [job_point_offsets[j], [j+1]) of points_xy
 * and polygons [job_polygon_offsets[j], [j+1]) of the GeoArrow polygon arrays),
 * job_bbox (4 per job: x_min, x_max, y_min, y_max), job_scale, job_max_depth,
 * job_max_size. Device inputs: points_xy (interleaved, 2 * num_points doubles)
 * and the polygons.
 * Outputs (device): point_indices (num_points; job j's point_indices at
 * [job_point_offsets[j], ...), values job-local), polygon_index and point_index
 * (*num_hits; job j's hits at [job_hit_offsets[j], job_hit_offsets[j+1]),
 * job-local, in the single-job order). job_hit_offsets is a HOST array of
 * num_jobs + 1 entries filled here.
 * Errors: the preconditions of the four single-job calls, per job.
 */
QPIP_API int qpip_batched_quadtree_point_in_polygon(int64_t num_jobs,
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
                                                    void* stream,
                                                    uint32_t** point_indices,
                                                    uint32_t** polygon_index,
                                                    uint32_t** point_index,
                                                    int64_t* job_hit_offsets,
                                                    int64_t* num_hits);

/*
 * Batched quadtree point-in-polygon as one captured HIP graph (no counterpart in
 * the reference). A plan fixes the jobs (the arguments of
 * qpip_batched_quadtree_point_in_polygon), the polygons and the ADDRESS of the
 * point buffer; each qpip_batch_plan_launch replays the whole pipeline as one
 * graph launch over the points the buffer then holds (rewrite them between
 * launches, keeping the job layout), with no host synchronisation and no
 * allocation. Per job the results equal qpip_batched_quadtree_point_in_polygon's
 * on the same points.
 *   create:  sizes the intermediate buffers by running the pipeline over the
 *            points in the buffer now (capacities = those counts x headroom,
 *            headroom in [1, 64]), then captures it. The polygons and the point
 *            buffer must outlive the plan.
 *   launch:  enqueues one replay on `stream`. Launches of one plan are
 *            serialised (a launch on another stream waits for the previous one).
 *   result:  waits for the last launch and returns its outputs:
 *            point_indices (num_points), polygon_index / point_index (*num_hits),
 *            job_hit_offsets (HOST, num_jobs + 1), as in the batched call. With
 *            a resource `mr` the columns are copied into memory allocated from
 *            it; with mr == NULL they are the plan's own buffers (valid until its
 *            next launch or destroy; do not qpip_free them).
 *            QPIP_ERROR_CAPACITY when the launch produced more leaf pairs, PIP
 *            chunks or hits than the plan holds (no output was written out of
 *            bounds; create the plan again from these points).
 */
typedef struct qpip_batch_plan qpip_batch_plan;
QPIP_API int qpip_batch_plan_create(int64_t num_jobs,
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
                                    void* stream,
                                    qpip_batch_plan** plan);
QPIP_API int qpip_batch_plan_launch(qpip_batch_plan* plan, void* stream);
QPIP_API int qpip_batch_plan_result(qpip_batch_plan* plan,
                                    const qpip_resource* mr,
                                    void* stream,
                                    uint32_t** point_indices,
                                    uint32_t** polygon_index,
                                    uint32_t** point_index,
                                    int64_t* job_hit_offsets,
                                    int64_t* num_hits);
QPIP_API int qpip_batch_plan_destroy(qpip_batch_plan* plan);

/*
 * linestring_bounding_boxes
 * Replaces cuspatial::linestring_bounding_boxes
 *   (cpp/include/cuspatial/bounding_boxes.hpp:52-57,
 *    cpp/src/bounding_boxes/linestring_bounding_boxes.cu:127-154;
 *    algorithm detail/bounding_boxes.cuh:97-131).
 * Box of each linestring's vertices, each vertex expanded by expansion_radius.
 * Outputs minx, miny, maxx, maxy (FLOAT64, *num_boxes = num_linestring_offsets - 1).
 * Errors: "x and y must be the same size", "expansion radius must be greater or
 * equal than 0", "all linestrings must have at least 2 vertices".
 */
QPIP_API int qpip_linestring_bounding_boxes(const int32_t* linestring_offsets,
                                            int64_t num_linestring_offsets,
                                            const double* x,
                                            const double* y,
                                            int64_t num_x,
                                            int64_t num_y,
                                            double expansion_radius,
                                            const qpip_resource* mr,
                                            void* stream,
                                            double** minx,
                                            double** miny,
                                            double** maxx,
                                            double** maxy,
                                            int64_t* num_boxes);

/*
 * quadtree_point_to_nearest_linestring
 * Replaces cuspatial::quadtree_point_to_nearest_linestring
 *   (cpp/include/cuspatial/spatial_join.hpp:166-175,
 *    cpp/src/join/quadtree_point_to_nearest_linestring.cu:128-189;
 *    algorithm detail/join/quadtree_point_to_nearest_linestring.cuh:143-314).
 * Inputs: the (linestring_index, quad_index) pairs of
 * qpip_join_quadtree_and_bounding_boxes over linestring boxes, the quadtree,
 * point_indices and x, y, and the linestrings (offsets, x, y; one linestring per
 * multilinestring). Outputs (num_point_indices rows): point_index (UINT32, the
 * row's position in point_indices), linestring_index (UINT32) and distance
 * (FLOAT64) of the nearest linestring among those paired with the point's leaf.
 * Rows of points in leaves with no pair: (p, 0xffffffff, 0.0) (the reference
 * leaves the index columns of these rows uninitialized).
 * Errors: "number of points must be the same for both x and y columns",
 * "numbers of vertices must be the same for both x and y columns".
 */
QPIP_API int qpip_quadtree_point_to_nearest_linestring(const uint32_t* linestring_index,
                                                       const uint32_t* quad_index,
                                                       int64_t num_pairs,
                                                       const uint32_t* key,
                                                       const uint8_t* level,
                                                       const uint8_t* is_internal_node,
                                                       const uint32_t* length,
                                                       const uint32_t* offset,
                                                       int64_t num_nodes,
                                                       const uint32_t* point_indices,
                                                       int64_t num_point_indices,
                                                       const double* x,
                                                       const double* y,
                                                       int64_t num_x,
                                                       int64_t num_y,
                                                       const uint32_t* linestring_offsets,
                                                       int64_t num_linestring_offsets,
                                                       const double* linestring_x,
                                                       const double* linestring_y,
                                                       int64_t num_linestring_x,
                                                       int64_t num_linestring_y,
                                                       const qpip_resource* mr,
                                                       void* stream,
                                                       uint32_t** point_index,
                                                       uint32_t** linestring_index_out,
                                                       double** distance,
                                                       int64_t* num_rows);

/*
 * Stable radix sort of (key, value) pairs on the low key_bits bits of the keys,
 * in place (device arrays of n uint32). Replaces thrust::stable_sort_by_key on
 * uint32 keys, the reference's sort primitive on this path
 * (detail/index/construction/phase_1.cuh:92,
 * detail/join/quadtree_bbox_filtering.cuh:161-165); the same kernels as the
 * quadtree's point sort. Pairs with equal keys keep their order. The host does
 * not wait.
 */
QPIP_API int qpip_radix_sort_pairs(uint32_t* keys,
                                   uint32_t* values,
                                   int64_t n,
                                   int32_t key_bits,
                                   void* stream);

/*
 * Stable compaction: the indices (and, when values_out is not NULL, the values)
 * of the entries of values[0..n) that are >= 0, in index order; *count of them.
 * indices_out / values_out need room for n entries. Replaces a stable
 * thrust::copy_if (the reference compacts its hits this way,
 * detail/join/quadtree_point_in_polygon.cuh:198-207). The host waits once, for
 * the count.
 */
QPIP_API int qpip_compact_nonnegative(const int32_t* values,
                                      int64_t n,
                                      uint32_t* indices_out,
                                      int32_t* values_out,
                                      int64_t* count,
                                      void* stream);

/*
 * Learned candidate ranking over a coarse grid (no counterpart in the
 * reference; an addition for point location over non-overlapping polygons).
 *
 * qpip_grid: a dim x dim grid of square cells of side `cell` from (x0, y0).
 * A point's cell is (floor((x - x0) * r), floor((y - y0) * r)) with r = 1 / cell
 * (one double division), each clamped to [0, dim - 1] (NaN -> 0), row-major
 * (iy * dim + ix). candidates
 * holds 16 polygon ids per cell (-1 padded) and counts the number of valid
 * slots; a count above 16 marks a cell whose points always take the exact
 * quadtree path. The candidate list of a cell must contain every polygon whose
 * bounding box reaches the cell (the locator's results are exact only then).
 *
 * qpip_ranker: the model, as one flat fp32 parameter array (params) and its bf16
 * copy (params_bf16, same element offsets): E[num_cells][64], W1[128][96],
 * b1[128], W2[128][128], b2[128], W3[16][128], b3[16] (qpip_ranker_num_params
 * floats). Features of a point: E[cell] and sin, cos of pi k u and pi k v for
 * k = 1..8, where (u, v) = ((x - (x0 + ix cell)) r, (y - (y0 + iy cell)) r) is
 * the point's position inside its cell in cell units (sine and cosine from the
 * hardware's v_sin_f32 / v_cos_f32).
 * logits = W3 relu(W2 relu(W1 x + b1) + b2) + b3 with bf16 operands and fp32
 * accumulation.
 */
typedef struct qpip_grid {
  double x0;
  double y0;
  double cell;
  int32_t dim;
  const int32_t* candidates; /* [dim * dim][16] */
  const uint8_t* counts;     /* [dim * dim] */
} qpip_grid;

typedef struct qpip_ranker {
  const float* params;
  const uint16_t* params_bf16;
  int64_t num_cells; /* must equal dim * dim of the grid it is used with */
} qpip_ranker;

/* Number of floats in the flat parameter array for num_cells embeddings. */
QPIP_API int64_t qpip_ranker_num_params(int64_t num_cells);

/* Feature rows (bf16 bits, [num_points][96]) and grid cells of points_xy. */
QPIP_API int qpip_ranker_features(const qpip_ranker* model,
                                  const qpip_grid* grid,
                                  const double* points_xy,
                                  int64_t num_points,
                                  uint16_t* features,
                                  uint32_t* cells,
                                  void* stream);

/* logits[num_points][16] of the model for points_xy (device arrays). */
QPIP_API int qpip_ranker_logits(const qpip_ranker* model,
                                const qpip_grid* grid,
                                const double* points_xy,
                                int64_t num_points,
                                float* logits,
                                void* stream);

/* Gradient of the mean masked softmax cross-entropy over a batch:
 * labels[i] is the slot (0..15) of point i's polygon in its cell's candidate
 * list, or -1 to leave the point out; the loss is summed over labelled points
 * and divided by num_points. grad (qpip_ranker_num_params floats) is
 * overwritten. stats (3 floats, device) accumulates: loss sum, top-1 correct,
 * labelled points. */
QPIP_API int qpip_ranker_gradients(const qpip_ranker* model,
                                   const qpip_grid* grid,
                                   const double* points_xy,
                                   const int32_t* labels,
                                   int64_t num_points,
                                   float* grad,
                                   float* stats,
                                   void* stream);

/* One Adam step (step counts from 1) on params; refreshes params_bf16. */
QPIP_API int qpip_ranker_adam(float* params,
                              uint16_t* params_bf16,
                              const float* grad,
                              float* m,
                              float* v,
                              int64_t num_params,
                              float lr,
                              float beta1,
                              float beta2,
                              float eps,
                              int64_t step,
                              void* stream);

/* Locate points: for each point, test the candidates of its cell in the order
 * of the model's logits (highest first) with the quadtree path's exact FP64
 * point-in-polygon test, at most max_tests of them, stopping at the first
 * polygon that contains the point.
 * out_polygon[i]: the polygon (>= 0), -1 when no candidate contains the point
 * (all were tested), or -2 when the point must take the exact quadtree path:
 * its cell overflowed, its hit lies within the collinearity filter of an edge
 * (another polygon's test could also hold), or max_tests ran out.
 * out_tests[i]: tests run. out_flags[i]: bit 0 quadtree path, bit 1 the first
 * prediction was not the answer. All arrays are device arrays. */
QPIP_API int qpip_ranker_locate(const qpip_ranker* model,
                                const qpip_grid* grid,
                                const int32_t* poly_offsets,
                                int64_t num_poly_offsets,
                                const int32_t* ring_offsets,
                                int64_t num_ring_offsets,
                                const double* poly_x,
                                const double* poly_y,
                                int64_t num_vertices,
                                const double* points_xy,
                                int64_t num_points,
                                int32_t max_tests,
                                int32_t* out_polygon,
                                uint8_t* out_tests,
                                uint8_t* out_flags,
                                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUSPATIAL_AMD_H */
