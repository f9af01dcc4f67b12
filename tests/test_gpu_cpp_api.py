"""The C++ interface (include/cuspatial_amd.hpp) driven by a compiled C++ program.

tests/cpp/test_cpp_api.cpp runs quadtree_on_points -> polygon_bounding_boxes ->
join_quadtree_and_bounding_boxes -> quadtree_point_in_polygon and
linestring_bounding_boxes -> join -> quadtree_point_to_nearest_linestring on the
GPU through the header, the same pipeline with a caller memory resource (the
reference's `mr` argument) on a created stream, then a call with mismatched x/y
that must throw cuspatial_amd::logic_error. Bar: every output column
bit-identical to the oracle; with the resource, every output column allocated
from it, equal to the pool-backed one, and released through it.
"""

import os
import re
import subprocess

import numpy as np
import pytest

from conftest import star_polygons

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "cpp", "_build", "test_cpp_api")


def _read(d, name, dtype):
    return np.fromfile(os.path.join(d, "out_" + name), dtype=dtype)


@pytest.mark.parametrize("seed,n,npoly,d,ms", [
    (11, 50000, 120, 6, 64),
    (12, 3000, 15, 4, 500),
])
def test_cpp_program_matches_oracle(tmp_path, oracle_mod, seed, n, npoly, d, ms):
    assert os.path.exists(EXE), "C++ test program not built (run __graft_entry__.build())"
    rng = np.random.default_rng(seed)
    ext = 100.0
    x = rng.uniform(0, ext, n)
    y = rng.uniform(0, ext, n)
    part, ring, px, py = star_polygons(rng, npoly, ext, 2.0, 9.0, holes=0.3)
    lo = np.array([0, 3, 5, 9, 12], np.int32)
    lx = rng.uniform(0, ext, 12)
    ly = rng.uniform(0, ext, 12)
    scale = ext / (1 << d)
    radius = 3.0
    prm = np.array([0.0, ext, 0.0, ext, scale, d, ms, radius], np.float64)
    for name, a in (("params.f64", prm), ("x.f64", x), ("y.f64", y), ("part.i32", part),
                    ("ring.i32", ring), ("px.f64", px), ("py.f64", py), ("lo.i32", lo),
                    ("lx.f64", lx), ("ly.f64", ly)):
        a.tofile(str(tmp_path / name))
    print(f"C++ program: seed {seed}, {n} points, {npoly} polygons", flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the reference's message format: "cuSpatial failure at: <file>:<line>: <reason>"
    assert "logic_error: cuSpatial failure at: " in r.stdout, r.stdout
    assert "x and y columns must have the same length" in r.stdout, r.stdout
    # with a caller resource (the reference's mr argument) every output column
    # comes from it, equals the pool-backed column, and goes back through it
    assert "resource outputs: 19 of 19 columns from the caller resource" in r.stdout, r.stdout
    assert "resource outputs equal the pool outputs: yes" in r.stdout, r.stdout
    m = re.search(r"resource after release: (\d+) allocations, (\d+) deallocations, "
                  r"(\d+) live, (\d+) bad frees", r.stdout)
    assert m, r.stdout
    allocs, deallocs, live, bad = map(int, m.groups())
    assert allocs >= 19 and allocs == deallocs and live == 0 and bad == 0, r.stdout

    opi, okey, olev, oint, olen, ooff = oracle_mod.quadtree_on_points(
        x, y, 0.0, ext, 0.0, ext, scale, d, ms)
    pi = _read(tmp_path, "point_indices", np.uint32)
    assert np.array_equal(pi, opi)
    assert np.array_equal(_read(tmp_path, "key", np.uint32), okey)
    assert np.array_equal(_read(tmp_path, "level", np.uint8), olev)
    assert np.array_equal(_read(tmp_path, "is_internal_node", np.uint8), oint.astype(np.uint8))
    assert np.array_equal(_read(tmp_path, "length", np.uint32), olen)
    assert np.array_equal(_read(tmp_path, "offset", np.uint32), ooff)
    tree = (okey, olev, oint, olen, ooff)

    boxes = oracle_mod.polygon_bounding_boxes(part, ring, px, py)
    opairs = oracle_mod.join_quadtree_and_bounding_boxes(tree, boxes, 0.0, ext, 0.0, ext, scale, d)
    assert np.array_equal(_read(tmp_path, "bbox_offset", np.uint32), opairs[0])
    assert np.array_equal(_read(tmp_path, "quad_offset", np.uint32), opairs[1])
    opoly, opt = oracle_mod.quadtree_point_in_polygon(opairs, tree, opi, x, y, part, ring, px, py)
    assert len(opoly) > 0
    assert np.array_equal(_read(tmp_path, "polygon_index", np.uint32), opoly)
    assert np.array_equal(_read(tmp_path, "point_index", np.uint32), opt)

    lboxes = oracle_mod.linestring_bounding_boxes(lo, lx, ly, radius)
    lpairs = oracle_mod.join_quadtree_and_bounding_boxes(tree, lboxes, 0.0, ext, 0.0, ext, scale,
                                                         d)
    npt, nls, ndist = oracle_mod.quadtree_point_to_nearest_linestring(
        lpairs, tree, opi, x, y, lo, lx, ly)
    assert np.array_equal(_read(tmp_path, "nearest_point_index", np.uint32), npt)
    assert np.array_equal(_read(tmp_path, "nearest_linestring_index", np.uint32), nls)
    assert np.array_equal(_read(tmp_path, "nearest_distance", np.float64).view(np.int64),
                          ndist.view(np.int64))
