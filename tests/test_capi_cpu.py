"""The C ABI without a GPU: the library loads, exports every symbol the header
declares, and reproduces the reference's precondition errors and empty-input
behaviour (both are decided before any device work)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "cuspatial_amd.h")


@pytest.fixture(scope="module")
def capi():
    from cuspatial_amd import _capi

    if not os.path.exists(_capi.LIB_PATH):
        from cuspatial_amd import _build

        _build.build()
    return _capi


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qpip_[a-z_]+)\s*\(", text)))


def test_exports_every_declared_symbol(capi):
    L = capi.lib()
    names = declared_symbols()
    assert len(names) == 35
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTED)
    assert "gfx950" in capi.version()


def test_exports_nothing_else(capi):
    """The library is built with hidden visibility: its dynamic symbol table
    holds exactly the declared entry points."""
    import shutil
    import subprocess

    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    text_syms = {ln.split()[-1] for ln in out.splitlines() if len(ln.split()) == 3 and
                 ln.split()[1] == "T"}
    assert text_syms == set(declared_symbols())


def _outs(n):
    return [ctypes.c_void_p() for _ in range(n)]


def _qt(capi, nx, ny, **kw):
    L = capi.lib()
    outs = _outs(6)
    m = ctypes.c_int64(-1)
    rc = L.qpip_quadtree_on_points(None, None, nx, ny, kw.get("x_min", 0.0), 1.0, 0.0, 1.0, 1.0,
                                   ctypes.c_int8(3), ctypes.c_int32(4), None, None,
                                   *[ctypes.byref(o) for o in outs], ctypes.byref(m))
    return rc, m.value, outs


def test_quadtree_size_mismatch(capi):
    rc, _, _ = _qt(capi, 3, 4)
    assert rc == capi.QPIP_ERROR_LOGIC
    msg = capi.lib().qpip_last_error_message().decode()
    assert msg.startswith("cuSpatial failure at: ")
    assert msg.endswith("x and y columns must have the same length")
    with pytest.raises(RuntimeError, match="x and y columns must have the same length"):
        capi.check(rc)


def test_quadtree_empty(capi):
    rc, m, outs = _qt(capi, 0, 0)
    assert rc == capi.QPIP_SUCCESS and m == 0
    assert all(o.value is None for o in outs)


def _join(capi, scale=1.0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, depth=3, m=0, b=0):
    L = capi.lib()
    outs = _outs(2)
    k = ctypes.c_int64(-1)
    rc = L.qpip_join_quadtree_and_bounding_boxes(
        None, None, None, None, None, m, None, None, None, None, b, xmin, xmax, ymin, ymax, scale,
        ctypes.c_int8(depth), None, None, *[ctypes.byref(o) for o in outs], ctypes.byref(k))
    return rc, k.value


@pytest.mark.parametrize("kw,msg", [
    (dict(scale=0.0), "scale must be positive"),
    (dict(scale=-1.0), "scale must be positive"),
    (dict(xmin=2.0), "invalid bounding box (x_min, x_max, y_min, y_max)"),
    (dict(ymax=0.0), "invalid bounding box (x_min, x_max, y_min, y_max)"),
    (dict(depth=0), "maximum depth must be positive and less than 16"),
    (dict(depth=16), "maximum depth must be positive and less than 16"),
])
def test_join_preconditions(capi, kw, msg):
    # quadtree_bbox_filtering.cu:102-106, in the reference's order
    rc, _ = _join(capi, **kw)
    assert rc == capi.QPIP_ERROR_LOGIC
    assert capi.lib().qpip_last_error_message().decode().endswith(msg)


def test_join_empty(capi):
    assert _join(capi, m=0, b=5) == (capi.QPIP_SUCCESS, 0)
    assert _join(capi, m=5, b=0) == (capi.QPIP_SUCCESS, 0)


def _pip(capi, pairs=0, nodes=0, npi=0, nx=0, ny=0, npo=0, nvx=0, nvy=0):
    L = capi.lib()
    outs = _outs(2)
    h = ctypes.c_int64(-1)
    rc = L.qpip_quadtree_point_in_polygon(
        None, None, pairs, None, None, None, None, None, nodes, None, npi, None, None, nx, ny,
        None, npo, None, 0, None, None, nvx, nvy, None, None, *[ctypes.byref(o) for o in outs],
        ctypes.byref(h))
    return rc, h.value


def test_pip_preconditions(capi):
    # quadtree_point_in_polygon.cu:158-161
    rc, _ = _pip(capi, npi=3, nx=3, ny=4)
    assert rc == capi.QPIP_ERROR_LOGIC
    assert capi.lib().qpip_last_error_message().decode().endswith(
        "number of points must be the same for both x and y columns")
    rc, _ = _pip(capi, npi=2, nx=3, ny=3)
    assert rc == capi.QPIP_ERROR_LOGIC
    rc, _ = _pip(capi, nvx=5, nvy=4)
    assert rc == capi.QPIP_ERROR_LOGIC
    assert capi.lib().qpip_last_error_message().decode().endswith(
        "numbers of vertices must be the same for both x and y columns")


@pytest.mark.parametrize("kw", [dict(pairs=0, nodes=3, npi=3, nx=3, ny=3, npo=2),
                                dict(pairs=3, nodes=0, npi=3, nx=3, ny=3, npo=2),
                                dict(pairs=3, nodes=3, npi=0, nx=0, ny=0, npo=2),
                                dict(pairs=3, nodes=3, npi=3, nx=3, ny=3, npo=0)])
def test_pip_empty(capi, kw):
    # quadtree_point_in_polygon.cu:171-178
    assert _pip(capi, **kw) == (capi.QPIP_SUCCESS, 0)


def test_bbox_preconditions_and_empty(capi):
    L = capi.lib()
    outs = _outs(4)
    nb = ctypes.c_int64(-1)
    rc = L.qpip_polygon_bounding_boxes(None, 3, None, 3, None, None, 4, 5, 0.0, None, None,
                                       *[ctypes.byref(o) for o in outs], ctypes.byref(nb))
    assert rc == capi.QPIP_ERROR_LOGIC
    assert capi.lib().qpip_last_error_message().decode().endswith("x and y must be the same size")
    for npo in (0, 1):
        rc = L.qpip_polygon_bounding_boxes(None, npo, None, 3, None, None, 4, 4, 0.0, None, None,
                                           *[ctypes.byref(o) for o in outs], ctypes.byref(nb))
        assert rc == capi.QPIP_SUCCESS and nb.value == 0
    rc = L.qpip_polygon_bounding_boxes(None, 3, None, 0, None, None, 4, 4, 0.0, None, None,
                                       *[ctypes.byref(o) for o in outs], ctypes.byref(nb))
    assert rc == capi.QPIP_ERROR_LOGIC
    assert capi.lib().qpip_last_error_message().decode().endswith(
        "Polygon ring offsets must contain at least one (1) value")


def test_python_scale_clamp_warning():
    """indexing.py:172-177: the warning text the reference emits."""
    import warnings

    import cuspatial_amd.spatial as sp

    x_min, x_max, y_min, y_max, scale, d = 0.0, 8.0, 0.0, 8.0, 0.01, 3
    min_scale = max(x_max - x_min, y_max - y_min) / ((1 << d) + 2)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        try:
            sp.join_quadtree_and_bounding_boxes(sp.Table(), [], x_min, x_max, y_min, y_max,
                                                scale, d)
        except RuntimeError:
            pass
    assert any(str(m.message) == "scale {} is less than required minimum scale {}. Clamping to "
               "minimum scale".format(scale, min_scale) for m in w)


def _brute(capi, pairwise, nx=0, ny=0, npo=1, nvx=0, nvy=0):
    L = capi.lib()
    out = ctypes.c_void_p()
    n = ctypes.c_int64(-1)
    fn = L.qpip_pairwise_point_in_polygon if pairwise else L.qpip_point_in_polygon
    rc = fn(None, None, nx, ny, None, npo, None, 1, None, None, nvx, nvy, None, None,
            ctypes.byref(out), ctypes.byref(n))
    return rc, n.value, out.value


@pytest.mark.parametrize("pairwise", [False, True])
def test_brute_force_pip_preconditions(capi, pairwise):
    # point_in_polygon.cu:118-136 and detail/point_in_polygon.cuh:94-95, in order
    rc, _, _ = _brute(capi, pairwise, nx=3, ny=4)
    assert rc == capi.QPIP_ERROR_LOGIC
    assert capi.lib().qpip_last_error_message().decode().endswith(
        "All points must have both x and y values")
    rc, _, _ = _brute(capi, pairwise, nvx=5, nvy=4)
    assert rc == capi.QPIP_ERROR_LOGIC
    if pairwise:
        rc, _, _ = _brute(capi, True, nx=2, ny=2, npo=2)
        assert rc == capi.QPIP_ERROR_LOGIC
        assert capi.lib().qpip_last_error_message().decode().endswith(
            "Must pass in the same number of points as polygons.")
    else:
        rc, _, _ = _brute(capi, False, npo=33)
        assert rc == capi.QPIP_ERROR_LOGIC
        assert capi.lib().qpip_last_error_message().decode().endswith(
            "Number of polygons cannot exceed 31")
        assert _brute(capi, False, npo=32)[0] == capi.QPIP_SUCCESS  # 31 polygons, no points


def test_brute_force_pip_empty(capi):
    assert _brute(capi, False) == (capi.QPIP_SUCCESS, 0, None)
    assert _brute(capi, True, npo=1) == (capi.QPIP_SUCCESS, 0, None)


def test_noalloc_pip_preconditions_and_limits(capi):
    """qpip_point_in_polygon_noalloc: the bitmask entry point's checks, then its
    scratch limits (262,144 vertices, 65,536 rings), all decided before any device
    work (the pointers below are never read)."""
    L = capi.lib()
    fake = ctypes.c_void_p(0x1000)

    def call(nx=4, ny=4, npo=2, nro=2, nvx=8, nvy=8, out=fake):
        return L.qpip_point_in_polygon_noalloc(fake, fake, nx, ny, fake, npo, fake, nro, fake,
                                               fake, nvx, nvy, None, out)

    msg = lambda: capi.lib().qpip_last_error_message().decode()  # noqa: E731
    assert call(nx=3) == capi.QPIP_ERROR_LOGIC
    assert msg().endswith("All points must have both x and y values")
    assert call(npo=33) == capi.QPIP_ERROR_LOGIC
    assert msg().endswith("Number of polygons cannot exceed 31")
    assert call(nx=0, ny=0) == capi.QPIP_SUCCESS
    assert call(out=None) == capi.QPIP_ERROR_LOGIC
    assert call(nvx=262145, nvy=262145) == capi.QPIP_ERROR_CAPACITY
    assert "262144" in msg() and "65536" in msg()
    assert call(nro=65538) == capi.QPIP_ERROR_CAPACITY


def test_batched_preconditions(capi):
    import numpy as np

    L = capi.lib()
    outs = _outs(3)
    hits = np.zeros(2, np.int64)
    h = ctypes.c_int64(-1)
    jpo = np.array([0, 5], np.int64)  # 5 points claimed, 3 given
    jpy = np.array([0, 0], np.int64)
    bb = np.array([0.0, 1.0, 0.0, 1.0])
    sc = np.array([1.0])
    md = np.array([3], np.int8)
    ms = np.array([4], np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = L.qpip_batched_quadtree_point_in_polygon(
        1, p(jpo), p(jpy), p(bb), p(sc), p(md), p(ms), None, 3, None, 1, None, 1, None, None,
        0, 0, None, None, *[ctypes.byref(o) for o in outs], p(hits), ctypes.byref(h))
    assert rc == capi.QPIP_ERROR_LOGIC
    assert "job point offsets must start at 0 and end at the number of points" in \
        L.qpip_last_error_message().decode()
    rc = L.qpip_batched_quadtree_point_in_polygon(
        0, None, None, None, None, None, None, None, 0, None, 0, None, 0, None, None, 0, 0,
        None, None, *[ctypes.byref(o) for o in outs], p(hits), ctypes.byref(h))
    assert rc == capi.QPIP_SUCCESS and h.value == 0 and hits[0] == 0


def test_ranker_preconditions(capi):
    """The ranker entry points check their arguments before touching the device."""
    from cuspatial_amd.ranker import qpip_grid, qpip_ranker

    L = capi.lib()
    assert L.qpip_ranker_num_params(0) == 96 * 128 + 128 + 128 * 128 + 128 + 16 * 128 + 16
    assert L.qpip_ranker_num_params(10) == L.qpip_ranker_num_params(0) + 640
    assert L.qpip_ranker_num_params(-1) == -1
    g = qpip_grid(0.0, 0.0, 1.0, 4, None, None)
    bad_dim = qpip_grid(0.0, 0.0, 1.0, 0, None, None)
    bad_cell = qpip_grid(0.0, 0.0, -1.0, 4, None, None)
    m = qpip_ranker(8, 8, 16)  # placeholder addresses: never dereferenced by the checks
    m_wrong = qpip_ranker(8, 8, 15)
    a = ctypes.addressof
    for grid, model, msg in ((bad_dim, m, "grid dimension must be in [1, 65536]"),
                             (bad_cell, m, "grid cell size must be positive"),
                             (g, m_wrong, "embedding count must equal the grid's cell count")):
        rc = L.qpip_ranker_logits(a(model), a(grid), None, 10, None, None)
        assert rc == capi.QPIP_ERROR_LOGIC and msg in L.qpip_last_error_message().decode()
    rc = L.qpip_ranker_locate(a(m), a(g), None, 1, None, 1, None, None, 0, None, 10, 0, None,
                              None, None, None)
    assert rc == capi.QPIP_ERROR_LOGIC
    assert "max_tests must be at least 1" in L.qpip_last_error_message().decode()
    rc = L.qpip_ranker_gradients(a(m), a(g), None, None, 0, None, None, None)
    assert rc == capi.QPIP_ERROR_LOGIC
    assert "at least one point" in L.qpip_last_error_message().decode()
    rc = L.qpip_ranker_adam(None, None, None, None, None, 4, 1e-3, 0.9, 0.999, 1e-8, 0, None)
    assert rc == capi.QPIP_ERROR_LOGIC and "Adam step" in L.qpip_last_error_message().decode()
    # empty inputs return without work
    assert L.qpip_ranker_logits(a(m), a(g), None, 0, None, None) == capi.QPIP_SUCCESS
    assert L.qpip_ranker_locate(a(m), a(g), None, 1, None, 1, None, None, 0, None, 0, 16, None,
                                None, None, None) == capi.QPIP_SUCCESS


def test_linestring_preconditions(capi):
    import numpy as np

    L = capi.lib()
    outs = _outs(4)
    n = ctypes.c_int64(-1)
    offs = np.array([0, 2, 3], np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = L.qpip_linestring_bounding_boxes(p(offs), 3, None, None, 3, 2, 0.0, None, None,
                                          *[ctypes.byref(o) for o in outs], ctypes.byref(n))
    assert rc == capi.QPIP_ERROR_LOGIC and "x and y must be the same size" in \
        L.qpip_last_error_message().decode()
    rc = L.qpip_linestring_bounding_boxes(p(offs), 3, None, None, 3, 3, -0.5, None, None,
                                          *[ctypes.byref(o) for o in outs], ctypes.byref(n))
    assert rc == capi.QPIP_ERROR_LOGIC and "expansion radius" in \
        L.qpip_last_error_message().decode()
    rc = L.qpip_linestring_bounding_boxes(p(offs), 3, None, None, 3, 3, 0.0, None, None,
                                          *[ctypes.byref(o) for o in outs], ctypes.byref(n))
    assert rc == capi.QPIP_ERROR_LOGIC and "at least 2 vertices" in \
        L.qpip_last_error_message().decode()
    rc = L.qpip_linestring_bounding_boxes(None, 0, None, None, 0, 0, 0.0, None, None,
                                          *[ctypes.byref(o) for o in outs], ctypes.byref(n))
    assert rc == capi.QPIP_SUCCESS and n.value == 0
    o3 = _outs(3)
    rc = L.qpip_quadtree_point_to_nearest_linestring(
        None, None, 1, None, None, None, None, None, 1, None, 5, None, None, 5, 4, None, 2, None,
        None, 3, 3, None, None, *[ctypes.byref(o) for o in o3], ctypes.byref(n))
    assert rc == capi.QPIP_ERROR_LOGIC and \
        "number of points must be the same for both x and y columns" in \
        L.qpip_last_error_message().decode()
    rc = L.qpip_quadtree_point_to_nearest_linestring(
        None, None, 0, None, None, None, None, None, 0, None, 0, None, None, 0, 0, None, 0, None,
        None, 0, 0, None, None, *[ctypes.byref(o) for o in o3], ctypes.byref(n))
    assert rc == capi.QPIP_SUCCESS and n.value == 0


def test_tuning_parameters(capi):
    """The autotuner's knobs: defaults, range checks, unknown names (host only)."""
    from cuspatial_amd import autotune

    assert autotune.get("pip_min_pairs_per_wave") == 2
    assert autotune.get("pip_pairs_shift") == 24
    assert autotune.get("pip_waves_per_block") == 2
    assert autotune.get("pip_stage_edges") == 64
    assert autotune.get("pip_xcd_run") == 16
    assert autotune.get("pip_box_filter") == 1
    assert autotune.get("pip_polygon_group") == 0
    with pytest.raises(RuntimeError, match="out of range"):
        autotune.set("pip_box_filter", 2)
    with pytest.raises(RuntimeError, match="out of range"):
        autotune.set("pip_stage_edges", 129)
    autotune.set("pip_min_pairs_per_wave", 64)
    assert autotune.current()["pip_min_pairs_per_wave"] == 64
    autotune.apply({"pip_min_pairs_per_wave": 2})
    with pytest.raises(RuntimeError, match="out of range"):
        autotune.set("pip_pairs_shift", 99)
    with pytest.raises(RuntimeError, match="unknown tuning parameter"):
        autotune.set("no_such_knob", 1)
    assert autotune.get("pip_pairs_shift") == 24


def test_batch_plan_preconditions(capi):
    """Plan creation checks its arguments before any device work; a capacity
    overflow has its own status code and exception."""
    L = capi.lib()
    plan = ctypes.c_void_p()
    jpo = (ctypes.c_int64 * 2)(0, 0)
    jpy = (ctypes.c_int64 * 2)(0, 0)
    bb = (ctypes.c_double * 4)(0.0, 1.0, 0.0, 1.0)
    sc = (ctypes.c_double * 1)(0.1)
    md = (ctypes.c_int8 * 1)(3)
    ms = (ctypes.c_int32 * 1)(8)

    def create(num_jobs, headroom):
        return L.qpip_batch_plan_create(num_jobs, jpo, jpy, bb, sc, md, ms, None, 0, None, 1,
                                        None, 1, None, None, 0, 0, headroom, None,
                                        ctypes.byref(plan))

    assert create(0, 1.25) == capi.QPIP_ERROR_LOGIC
    assert "a plan needs at least one job" in capi.lib().qpip_last_error_message().decode()
    assert create(1, 0.5) == capi.QPIP_ERROR_LOGIC
    assert "headroom must be in [1, 64]" in capi.lib().qpip_last_error_message().decode()
    assert plan.value is None
    assert L.qpip_batch_plan_launch(None, None) == capi.QPIP_ERROR_LOGIC
    assert "plan must not be null" in capi.lib().qpip_last_error_message().decode()
    assert L.qpip_batch_plan_destroy(None) == capi.QPIP_SUCCESS
    with pytest.raises(capi.CapacityError):
        capi.check(capi.QPIP_ERROR_CAPACITY)
