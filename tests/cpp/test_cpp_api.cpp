// Drives the C++ interface (include/cuspatial_amd.hpp) end to end on the GPU:
// quadtree -> polygon boxes -> join -> point-in-polygon, and linestring boxes ->
// join -> nearest linestring, reading its inputs from raw binary files written by
// tests/test_gpu_cpp_api.py and writing every output column back for comparison
// with the oracle. Usage: test_cpp_api <dir>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <type_traits>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/cuspatial_amd.hpp"

namespace {

std::string g_dir;

template <class T>
std::vector<T> read_bin(char const* name)
{
  std::ifstream f(g_dir + "/" + name, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error(std::string("cannot open ") + name);
  auto const bytes = static_cast<size_t>(f.tellg());
  std::vector<T> v(bytes / sizeof(T));
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(bytes));
  return v;
}

template <class T>
T* to_device(std::vector<T> const& v)
{
  T* p = nullptr;
  if (v.empty()) return nullptr;
  if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) throw std::bad_alloc();
  if (hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("hipMemcpy to device failed");
  return p;
}

template <class T>
void write_col(char const* name, cuspatial_amd::column<T> const& c)
{
  std::vector<T> h(static_cast<size_t>(c.size()));
  if (c.size() > 0 &&
      hipMemcpy(h.data(), c.data(), h.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("hipMemcpy to host failed");
  std::ofstream f(g_dir + "/out_" + name, std::ios::binary);
  f.write(reinterpret_cast<char const*>(h.data()), static_cast<std::streamsize>(h.size() * sizeof(T)));
}

// A caller memory resource that counts what it hands out (the role of an
// rmm::device_async_resource_ref in the reference's API).
struct CountingResource {
  std::map<void*, size_t> live;
  size_t allocations   = 0;
  size_t deallocations = 0;
  size_t bad_frees     = 0;
  static void* allocate(void* ctx, size_t bytes, void*)
  {
    auto* r  = static_cast<CountingResource*>(ctx);
    void* p  = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    r->live[p] = bytes;
    ++r->allocations;
    return p;
  }
  static void deallocate(void* ctx, void* p, size_t bytes, void* stream)
  {
    auto* r = static_cast<CountingResource*>(ctx);
    auto it = r->live.find(p);
    if (it == r->live.end() || it->second != bytes) {
      ++r->bad_frees;
      return;
    }
    r->live.erase(it);
    ++r->deallocations;
    (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
    (void)hipFree(p);
  }
  bool owns(void const* p) const { return live.count(const_cast<void*>(p)) != 0; }
};

}  // namespace

int main(int argc, char** argv)
{
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  try {
    namespace cs = cuspatial_amd;
    auto const prm = read_bin<double>("params.f64");  // x0 x1 y0 y1 scale depth max_size radius
    auto const x   = read_bin<double>("x.f64");
    auto const y   = read_bin<double>("y.f64");
    auto const part = read_bin<int32_t>("part.i32");
    auto const ring = read_bin<int32_t>("ring.i32");
    auto const px   = read_bin<double>("px.f64");
    auto const py   = read_bin<double>("py.f64");
    auto const lo   = read_bin<int32_t>("lo.i32");
    auto const lx   = read_bin<double>("lx.f64");
    auto const ly   = read_bin<double>("ly.f64");
    double *dx = to_device(x), *dy = to_device(y), *dpx = to_device(px), *dpy = to_device(py);
    double *dlx = to_device(lx), *dly = to_device(ly);
    int32_t *dpart = to_device(part), *dring = to_device(ring), *dlo = to_device(lo);
    int64_t const n = static_cast<int64_t>(x.size());
    auto const d = static_cast<int8_t>(prm[5]);

    auto [point_indices, tree] = cs::quadtree_on_points(dx, n, dy, n, prm[0], prm[1], prm[2],
                                                        prm[3], prm[4], d,
                                                        static_cast<int32_t>(prm[6]));
    auto boxes = cs::polygon_bounding_boxes(dpart, static_cast<int64_t>(part.size()), dring,
                                            static_cast<int64_t>(ring.size()), dpx,
                                            static_cast<int64_t>(px.size()), dpy,
                                            static_cast<int64_t>(py.size()));
    auto pairs = cs::join_quadtree_and_bounding_boxes(tree, boxes, prm[0], prm[1], prm[2], prm[3],
                                                      prm[4], d);
    auto hits  = cs::quadtree_point_in_polygon(
      pairs, tree, point_indices, dx, n, dy, n, dpart, static_cast<int64_t>(part.size()), dring,
      static_cast<int64_t>(ring.size()), dpx, static_cast<int64_t>(px.size()), dpy,
      static_cast<int64_t>(py.size()));
    auto lboxes = cs::linestring_bounding_boxes(dlo, static_cast<int64_t>(lo.size()), dlx,
                                                static_cast<int64_t>(lx.size()), dly,
                                                static_cast<int64_t>(ly.size()), prm[7]);
    auto lpairs = cs::join_quadtree_and_bounding_boxes(tree, lboxes, prm[0], prm[1], prm[2],
                                                       prm[3], prm[4], d);
    auto nearest = cs::quadtree_point_to_nearest_linestring(
      lpairs, tree, point_indices, dx, n, dy, n, reinterpret_cast<uint32_t const*>(dlo),
      static_cast<int64_t>(lo.size()), dlx, static_cast<int64_t>(lx.size()), dly,
      static_cast<int64_t>(ly.size()));

    // the same pipeline with every output from a caller resource
    CountingResource counter;
    qpip_resource const res{&CountingResource::allocate, &CountingResource::deallocate, &counter};
    hipStream_t stream = nullptr;
    if (hipStreamCreate(&stream) != hipSuccess) throw std::runtime_error("hipStreamCreate failed");
    {
      auto [r_pi, r_tree] = cs::quadtree_on_points(dx, n, dy, n, prm[0], prm[1], prm[2], prm[3],
                                                   prm[4], d, static_cast<int32_t>(prm[6]),
                                                   stream, &res);
      auto r_boxes = cs::polygon_bounding_boxes(dpart, static_cast<int64_t>(part.size()), dring,
                                                static_cast<int64_t>(ring.size()), dpx,
                                                static_cast<int64_t>(px.size()), dpy,
                                                static_cast<int64_t>(py.size()), 0.0, stream, &res);
      auto r_pairs = cs::join_quadtree_and_bounding_boxes(r_tree, r_boxes, prm[0], prm[1], prm[2],
                                                          prm[3], prm[4], d, stream, &res);
      auto r_hits  = cs::quadtree_point_in_polygon(
        r_pairs, r_tree, r_pi, dx, n, dy, n, dpart, static_cast<int64_t>(part.size()), dring,
        static_cast<int64_t>(ring.size()), dpx, static_cast<int64_t>(px.size()), dpy,
        static_cast<int64_t>(py.size()), stream, &res);
      auto r_lboxes = cs::linestring_bounding_boxes(dlo, static_cast<int64_t>(lo.size()), dlx,
                                                    static_cast<int64_t>(lx.size()), dly,
                                                    static_cast<int64_t>(ly.size()), prm[7],
                                                    stream, &res);
      auto r_lpairs = cs::join_quadtree_and_bounding_boxes(r_tree, r_lboxes, prm[0], prm[1],
                                                           prm[2], prm[3], prm[4], d, stream, &res);
      auto r_near = cs::quadtree_point_to_nearest_linestring(
        r_lpairs, r_tree, r_pi, dx, n, dy, n, reinterpret_cast<uint32_t const*>(dlo),
        static_cast<int64_t>(lo.size()), dlx, static_cast<int64_t>(lx.size()), dly,
        static_cast<int64_t>(ly.size()), stream, &res);
      if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("stream failed");
      std::vector<void const*> outs = {
        r_pi.data(), r_tree.key.data(), r_tree.level.data(), r_tree.is_internal_node.data(),
        r_tree.length.data(), r_tree.offset.data(), r_boxes.minx.data(), r_boxes.miny.data(),
        r_boxes.maxx.data(), r_boxes.maxy.data(), r_pairs.bbox_offset.data(),
        r_pairs.quad_offset.data(), r_hits.polygon_index.data(), r_hits.point_index.data(),
        r_lboxes.minx.data(), r_lpairs.bbox_offset.data(), r_near.point_index.data(),
        r_near.linestring_index.data(), r_near.distance.data()};
      size_t from_resource = 0;
      for (void const* p : outs) from_resource += counter.owns(p) ? 1 : 0;
      std::printf("resource outputs: %zu of %zu columns from the caller resource, %zu live\n",
                  from_resource, outs.size(), counter.live.size());
      // same columns as the pool-backed run
      auto same = [](auto const& a, auto const& b) {
        using T = std::remove_pointer_t<decltype(a.data())>;
        if (a.size() != b.size()) return false;
        std::vector<T> ha(static_cast<size_t>(a.size())), hb(ha.size());
        if (a.size() == 0) return true;
        if (hipMemcpy(ha.data(), a.data(), ha.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hb.data(), b.data(), hb.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
          return false;
        return std::memcmp(ha.data(), hb.data(), ha.size() * sizeof(T)) == 0;
      };
      bool const equal = same(r_pi, point_indices) && same(r_tree.key, tree.key) &&
                         same(r_tree.offset, tree.offset) && same(r_pairs.quad_offset, pairs.quad_offset) &&
                         same(r_hits.polygon_index, hits.polygon_index) &&
                         same(r_hits.point_index, hits.point_index) &&
                         same(r_near.distance, nearest.distance);
      std::printf("resource outputs equal the pool outputs: %s\n", equal ? "yes" : "no");
    }
    std::printf("resource after release: %zu allocations, %zu deallocations, %zu live, %zu bad frees\n",
                counter.allocations, counter.deallocations, counter.live.size(), counter.bad_frees);
    (void)hipStreamDestroy(stream);

    write_col("point_indices", point_indices);
    write_col("key", tree.key);
    write_col("level", tree.level);
    write_col("is_internal_node", tree.is_internal_node);
    write_col("length", tree.length);
    write_col("offset", tree.offset);
    write_col("bbox_offset", pairs.bbox_offset);
    write_col("quad_offset", pairs.quad_offset);
    write_col("polygon_index", hits.polygon_index);
    write_col("point_index", hits.point_index);
    write_col("nearest_point_index", nearest.point_index);
    write_col("nearest_linestring_index", nearest.linestring_index);
    write_col("nearest_distance", nearest.distance);

    // a violated precondition surfaces as cuspatial_amd::logic_error with the
    // reference's reason (src/indexing/point_quadtree.cu:166)
    try {
      (void)cs::quadtree_on_points(dx, n, dy, n - 1, prm[0], prm[1], prm[2], prm[3], prm[4], d, 1);
      std::printf("no error for mismatched x and y\n");
      return 1;
    } catch (cs::logic_error const& e) {
      std::printf("logic_error: %s\n", e.what());
    }
    for (void* p : {static_cast<void*>(dx), static_cast<void*>(dy), static_cast<void*>(dpx),
                    static_cast<void*>(dpy), static_cast<void*>(dlx), static_cast<void*>(dly),
                    static_cast<void*>(dpart), static_cast<void*>(dring), static_cast<void*>(dlo)})
      if (p) (void)hipFree(p);
    std::printf("cpp api ok: %lld nodes, %lld pairs, %lld hits, %lld nearest rows\n",
                static_cast<long long>(tree.size()), static_cast<long long>(pairs.size()),
                static_cast<long long>(hits.size()), static_cast<long long>(nearest.size()));
    return 0;
  } catch (std::exception const& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
