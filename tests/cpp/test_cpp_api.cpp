// Drives the C++ interface (include/cuspatial_amd.hpp) end to end on the GPU:
// quadtree -> polygon boxes -> join -> point-in-polygon, and linestring boxes ->
// join -> nearest linestring, reading its inputs from raw binary files written by
// tests/test_gpu_cpp_api.py and writing every output column back for comparison
// with the oracle. Usage: test_cpp_api <dir>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
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
