"""Parity of the HIP path (through the C ABI) with the oracle.

Bar: bit-identical quadtree columns, point_indices, bounding boxes, join pairs
and point-in-polygon pairs (integer/index outputs and exact FP64 predicates).
"""

import ctypes

import numpy as np
import pytest

from conftest import load_golden, star_polygons

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cs():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd

    return cuspatial_amd


def np_(t):
    return t.cpu().numpy()


def i64(t):
    """uint32 device column -> int64 (through an int32 view: torch's uint32
    support on the GPU is limited to copies)."""
    import torch

    if t.dtype == torch.uint32:
        return t.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    return t.to(torch.int64)


def u32(t):
    import torch

    return t.to(torch.int32).view(torch.uint32)


def clamp_scale(x_min, x_max, y_min, y_max, scale, d):
    x_min, x_max = min(x_min, x_max), max(x_min, x_max)
    y_min, y_max = min(y_min, y_max), max(y_min, y_max)
    return max(scale, max(x_max - x_min, y_max - y_min) / ((1 << d) + 2))


def make_points(cs, x, y, layout):
    """Column layout (x, y arrays) or interleaved xy (the GeoSeries / vec_2d layout);
    "xy8" places the interleaved array 8 bytes off a 16-byte boundary."""
    import torch

    if layout == "columns":
        return cs.Points(x, y)
    xt = torch.as_tensor(np.asarray(x, np.float64)) if not isinstance(x, torch.Tensor) else x
    yt = torch.as_tensor(np.asarray(y, np.float64)) if not isinstance(y, torch.Tensor) else y
    xy = torch.stack([xt.to("cuda"), yt.to("cuda")], dim=1).reshape(-1)
    if layout == "xy8":
        buf = torch.empty(xy.numel() + 1, dtype=torch.float64, device="cuda")
        buf[1:] = xy
        xy = buf[1:]
        assert xy.data_ptr() % 16 == 8
    return cs.Points.from_points_xy(xy)


def run_gpu(cs, x, y, bbox, scale, d, max_size, polys=None, layout="columns"):
    pts = make_points(cs, x, y, layout)
    pi, tree = cs.quadtree_on_points(pts, *bbox, scale, d, max_size)
    out = {"pi": np_(pi)}
    out.update({k: np_(v) for k, v in tree.items()})
    if polys is not None and len(tree) > 0:
        P = cs.Polygons(*polys)
        bb = cs.polygon_bounding_boxes(P)
        out.update({k: np_(v) for k, v in bb.items()})
        pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *bbox, scale, d)
        out.update({k: np_(v) for k, v in pairs.items()})
        res = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, P)
        out.update({k: np_(v) for k, v in res.items()})
    return out


def run_oracle(oracle, x, y, bbox, scale, d, max_size, polys=None):
    pi, key, level, internal, length, offset = oracle.quadtree_on_points(
        x, y, *bbox, scale, d, max_size)
    out = dict(pi=pi, key=key, level=level, is_internal_node=internal, length=length,
               offset=offset)
    if polys is not None and len(key) > 0:
        part, ring, vx, vy = polys
        bb = oracle.polygon_bounding_boxes(part, ring, vx, vy)
        out.update(dict(zip(["minx", "miny", "maxx", "maxy"], bb)))
        # the Python layer orders the bbox and clamps the scale (join.py:144-156)
        s = clamp_scale(*bbox, scale, d)
        x0, x1, y0, y1 = bbox
        ob = (min(x0, x1), max(x0, x1), min(y0, y1), max(y0, y1))
        jb, jq = oracle.join_quadtree_and_bounding_boxes(
            (key, level, internal, length, offset), bb, *ob, s, d)
        out.update(bbox_offset=jb, quad_offset=jq)
        pp, pt = oracle.quadtree_point_in_polygon((jb, jq), (key, level, internal, length, offset),
                                                  pi, x, y, part, ring, vx, vy)
        out.update(polygon_index=pp, point_index=pt)
    return out


def assert_same(g, o):
    assert set(o) <= set(g), (sorted(g), sorted(o))
    for k in o:
        a, b = np.asarray(g[k]), np.asarray(o[k])
        if k in ("minx", "miny", "maxx", "maxy"):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)) or \
                np.array_equal(a, b, equal_nan=True), k
        else:
            assert a.shape == b.shape, (k, a.shape, b.shape)
            assert np.array_equal(a.astype(np.int64), b.astype(np.int64)), k


# ---------------------------------------------------------------------------
# golden vectors
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", load_golden("quadtree_cases.json"), ids=lambda c: c["name"])
def test_quadtree_golden(cs, case):
    p = np.array(case["points"], dtype=np.float64).reshape(-1, 2)
    v1, v2 = case["vertex_1"], case["vertex_2"]
    g = run_gpu(cs, p[:, 0], p[:, 1], (v1[0], v2[0], v1[1], v2[1]), case["scale"],
                case["max_depth"], case["max_size"])
    assert g["key"].tolist() == case["key"]
    assert g["level"].tolist() == case["level"]
    assert g["is_internal_node"].tolist() == case["is_internal_node"]
    assert g["length"].tolist() == case["length"]
    assert g["offset"].tolist() == case["offset"]
    assert len(g["pi"]) == len(p)


def test_join_pip_golden(cs):
    g = load_golden("join_pip_small.json")
    p = np.array(g["points"])
    v = np.array(g["vertices"])
    out = run_gpu(cs, p[:, 0], p[:, 1], (0.0, 8.0, 0.0, 8.0), g["scale"], g["max_depth"],
                  g["max_size"], (g["part_offsets"], g["ring_offsets"], v[:, 0], v[:, 1]))
    assert out["bbox_offset"].tolist() == g["expected_join_bbox_offset"]
    assert out["quad_offset"].tolist() == g["expected_join_quad_offset"]
    assert out["polygon_index"].tolist() == g["expected_polygon_index"]
    assert out["point_index"].tolist() == g["expected_point_index"]


@pytest.mark.parametrize("case", load_golden("pip_kats.json"), ids=lambda c: c["name"])
def test_pip_kats_through_quadtree(cs, oracle_mod, case):
    """The predicate KATs, run through the whole quadtree path."""
    pts = np.array(case["points"], dtype=np.float64).reshape(-1, 2)
    v = np.array(case["vertices"])
    bbox = (-4.0, 4.0, -4.0, 4.0)
    polys = (np.array(case["part_offsets"], np.int32), np.array(case["ring_offsets"], np.int32),
             v[:, 0], v[:, 1])
    g = run_gpu(cs, pts[:, 0], pts[:, 1], bbox, 1.0, 3, 1, polys)
    o = run_oracle(oracle_mod, pts[:, 0], pts[:, 1], bbox, 1.0, 3, 1, polys)
    assert_same(g, o)
    bits = np.zeros(len(pts), np.int64)
    for poly, pos in zip(g["polygon_index"], g["point_index"]):
        bits[g["pi"][pos]] |= 1 << int(poly)
    assert bits.tolist() == case["expected"]


# ---------------------------------------------------------------------------
# seeded random cases against the oracle
# ---------------------------------------------------------------------------
def scene(seed, n, ext=100.0, npoly=40, outside=0.05, dup=0.02):
    rng = np.random.default_rng(seed)
    k = int(n * 0.6)
    x = np.concatenate([rng.normal(ext * 0.3, ext * 0.08, k), rng.uniform(0, ext, n - k)])
    y = np.concatenate([rng.normal(ext * 0.6, ext * 0.05, k), rng.uniform(0, ext, n - k)])
    x = np.clip(x, 0, np.nextafter(ext, 0))
    y = np.clip(y, 0, np.nextafter(ext, 0))
    m = int(n * outside)
    if m:
        x[:m] = rng.uniform(-0.2 * ext, 1.2 * ext, m)
        y[:m] = rng.uniform(-0.2 * ext, 1.2 * ext, m)
    ndup = int(n * dup)
    if ndup and n > 2 * ndup:
        x[-ndup:] = x[m:m + ndup]
        y[-ndup:] = y[m:m + ndup]
    polys = star_polygons(rng, npoly, ext, ext * 0.01, ext * 0.08, 3, 40, holes=0.3)
    return x, y, polys


@pytest.mark.parametrize("seed,n,d,max_size", [
    (1, 1, 3, 1), (2, 2, 1, 1), (3, 100, 0, 1), (4, 300, 1, 5), (5, 1000, 2, 3),
    (6, 5000, 5, 12), (7, 20000, 7, 125), (8, 50000, 9, 30), (9, 100000, 12, 64),
    (10, 200000, 15, 256), (11, 100000, 15, 1), (12, 70000, 6, 100000), (13, 4097, 4, 16),
    (14, 65537, 8, 8), (15, 30000, 11, 2)])
@pytest.mark.parametrize("layout", ["columns", "xy"])
def test_random_parity(cs, oracle_mod, seed, n, d, max_size, layout):
    x, y, polys = scene(seed, n)
    bbox = (0.0, 100.0, 0.0, 100.0)
    scale = 100.0 / (1 << d)
    print(f"[parity] seed={seed} n={n} d={d} max_size={max_size} {layout}", flush=True)
    # the join requires 0 < max_depth < 16 (quadtree_bbox_filtering.cu:104-105)
    p = polys if d > 0 else None
    g = run_gpu(cs, x, y, bbox, scale, d, max_size, p, layout=layout)
    o = run_oracle(oracle_mod, x, y, bbox, scale, d, max_size, p)
    assert_same(g, o)
    if d == 0:
        with pytest.raises(RuntimeError, match="maximum depth must be positive"):
            run_gpu(cs, x, y, bbox, scale, d, max_size, polys)


@pytest.mark.parametrize("seed,d,max_size", [(21, 2, 1), (22, 3, 2), (23, 1, 1), (24, 6, 4),
                                             (25, 15, 3), (26, 1, 100)])
@pytest.mark.parametrize("layout", ["columns", "xy"])
def test_clamped_scale_and_max_edge(cs, oracle_mod, seed, d, max_size, layout):
    """Scale below extent / 2^d is clamped to extent / (2^d + 2); points exactly on
    the bbox maximum then reach cell coordinate >= 2^d and the root level holds
    more than one node (the reference's num_parent_nodes quirk)."""
    rng = np.random.default_rng(seed)
    n = 3000
    x = rng.uniform(-1000, 1000, n)
    y = rng.uniform(-1000, 1000, n)
    x[:40] = 1000.0
    y[20:60] = 1000.0
    x[60:70] = -1000.0
    bbox = (1000.0, -1000.0, -1000.0, 1000.0)  # unordered on purpose
    polys = star_polygons(rng, 20, 2000.0, 50.0, 300.0, holes=0.2)
    polys = (polys[0], polys[1], polys[2] - 1000.0, polys[3] - 1000.0)
    g = run_gpu(cs, x, y, bbox, -1.0, d, max_size, polys, layout=layout)
    o = run_oracle(oracle_mod, x, y, bbox, -1.0, d, max_size, polys)
    assert_same(g, o)


def test_all_points_outside(cs, oracle_mod):
    rng = np.random.default_rng(31)
    x = rng.uniform(200, 300, 5000)
    y = rng.uniform(-300, -200, 5000)
    polys = star_polygons(rng, 10, 100.0, 5.0, 20.0)
    g = run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100.0 / 64, 6, 10, polys)
    o = run_oracle(oracle_mod, x, y, (0.0, 100.0, 0.0, 100.0), 100.0 / 64, 6, 10, polys)
    assert_same(g, o)


def test_empty_inputs(cs):
    import torch

    e = torch.empty(0, dtype=torch.float64, device="cuda")
    pi, tree = cs.quadtree_on_points(cs.Points(e, e), 0, 1, 0, 1, 1.0, 1, 1)
    assert len(pi) == 0 and len(tree) == 0 and tree.columns == [
        "key", "level", "is_internal_node", "length", "offset"]
    bb = cs.polygon_bounding_boxes(cs.Polygons([0], [0], e, e))
    assert len(bb) == 0 and bb.columns == ["minx", "miny", "maxx", "maxy"]
    pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, 0, 1, 0, 1, 1.0, 1)
    assert len(pairs) == 0 and pairs.columns == ["bbox_offset", "quad_offset"]


def test_empty_polygon_bbox_semantics(cs, oracle_mod):
    """reduce_by_key emits one box per non-empty polygon, in order (an empty
    polygon shifts later boxes up); rows with no box are NaN here."""
    part = np.array([0, 1, 1, 2, 3], np.int32)  # polygon 1 has no ring
    ring = np.array([0, 4, 4, 8], np.int32)     # ring 1 (polygon 2) is empty
    vx = np.array([0.0, 1, 1, 0, 5, 6, 6, 5])
    vy = np.array([0.0, 0, 1, 1, 5, 5, 6, 6])
    bb = cs.polygon_bounding_boxes(cs.Polygons(part, ring, vx, vy))
    o = oracle_mod.polygon_bounding_boxes(part, ring, vx, vy)
    for k, ov in zip(["minx", "miny", "maxx", "maxy"], o):
        assert np.array_equal(np_(bb[k]), ov, equal_nan=True), (k, np_(bb[k]), ov)


# ---------------------------------------------------------------------------
# the two workloads
# ---------------------------------------------------------------------------
def test_uniform_stars_full_parity(cs, oracle_mod):
    from cuspatial_amd import datasets

    w = datasets.uniform_stars()
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    polys = (w.part_offset, w.ring_offset, w.px, w.py)
    g = run_gpu(cs, w.x, w.y, bbox, w.scale, w.max_depth, w.max_size, polys)
    o = run_oracle(oracle_mod, w.x, w.y, bbox, w.scale, w.max_depth, w.max_size, polys)
    assert_same(g, o)
    assert len(o["polygon_index"]) > 10000


@pytest.mark.parametrize("layout", ["columns", "xy", "xy8"])
def test_gaussian_holes_sample_parity(cs, oracle_mod, layout):
    from cuspatial_amd import datasets

    w = datasets.gaussian_holes(num_points=1_000_000)
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    polys = (w.part_offset, w.ring_offset, w.px, w.py)
    g = run_gpu(cs, w.x, w.y, bbox, w.scale, w.max_depth, w.max_size, polys, layout=layout)
    o = run_oracle(oracle_mod, w.x, w.y, bbox, w.scale, w.max_depth, w.max_size, polys)
    assert_same(g, o)
    assert len(o["polygon_index"]) > 1000


def _spread_t(v):
    import torch

    v = v.to(torch.int64) & 0xFFFF
    v = (v | (v << 8)) & 0x00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F
    v = (v | (v << 2)) & 0x33333333
    v = (v | (v << 1)) & 0x55555555
    return v


def test_gaussian_holes_full_size_properties(cs, oracle_mod):
    """At the full 1e8-point size: structural properties of the tree on the
    device, the join against the oracle on the GPU-built tree (exact), and the
    point-in-polygon against the oracle on a strided subset of the pairs (exact)."""
    import torch

    from cuspatial_amd import datasets

    w = datasets.gaussian_holes(device=torch.device("cuda"))
    n, d, ms = w.num_points, w.max_depth, w.max_size
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    pts = cs.Points(w.x, w.y)
    pi, tree = cs.quadtree_on_points(pts, *bbox, w.scale, d, ms)
    print(f"[full] n={n} nodes={len(tree)}", flush=True)
    # point_indices is a permutation
    assert torch.equal(torch.sort(i64(pi))[0], torch.arange(n, device="cuda"))
    key = i64(tree["key"])
    level = tree["level"].to(torch.int64)
    internal = tree["is_internal_node"]
    length = i64(tree["length"])
    offset = i64(tree["offset"])
    # leaves tile [0, n) in offset order
    leaf = ~internal
    lo, order = torch.sort(offset[leaf])
    ll = length[leaf][order]
    assert int(lo[0]) == 0 and int(ll.sum()) == n
    assert torch.equal(lo[1:], (lo + ll)[:-1])
    # every point of a leaf has the leaf's key prefix (keys recomputed here)
    xs, ys = w.x, w.y
    cx = torch.trunc((xs - w.x_min) / w.scale).to(torch.int64)
    cy = torch.trunc((ys - w.y_min) / w.scale).to(torch.int64)
    pkey = (_spread_t(cy) << 1) | _spread_t(cx)
    pos_leaf = torch.searchsorted(lo, torch.arange(n, device="cuda"), right=True) - 1
    leaf_key = key[leaf][order][pos_leaf]
    leaf_lvl = level[leaf][order][pos_leaf]
    pk = pkey[i64(pi)]
    assert torch.equal(pk >> (2 * (d - 1 - leaf_lvl)), leaf_key)
    # internal iff more than max_size points (level <= d - 2); children are contiguous
    assert bool((length[internal] >= 1).all()) and bool((length[internal] <= 4).all())
    assert bool((leaf_lvl <= d - 1).all())
    del cx, cy, pkey, pos_leaf, leaf_key, leaf_lvl, pk
    # join: oracle on the GPU-built tree
    P = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    bb = cs.polygon_bounding_boxes(P)
    pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *bbox, w.scale, d)
    t5 = tuple(np_(tree[k]) for k in ("key", "level", "is_internal_node", "length", "offset"))
    obb = oracle_mod.polygon_bounding_boxes(w.part_offset, w.ring_offset, w.px, w.py)
    for k, ov in zip(["minx", "miny", "maxx", "maxy"], obb):
        assert np.array_equal(np_(bb[k]), ov)
    jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(t5, obb, *bbox, w.scale, d)
    assert np.array_equal(np_(pairs["bbox_offset"]), jb)
    assert np.array_equal(np_(pairs["quad_offset"]), jq)
    print(f"[full] pairs={len(jb)}", flush=True)
    # PIP on every 37th pair (~3% of candidates): GPU vs oracle, same pair list
    sel = torch.arange(0, len(jb), 37, device="cuda")
    sub = cs.Table(bbox_offset=u32(i64(pairs["bbox_offset"])[sel]),
                   quad_offset=u32(i64(pairs["quad_offset"])[sel]))
    res = cs.quadtree_point_in_polygon(sub, tree, pi, pts, P)
    pp, pt = oracle_mod.quadtree_point_in_polygon(
        (jb[::37], jq[::37]), t5, np_(pi), np_(w.x), np_(w.y), w.part_offset, w.ring_offset,
        w.px, w.py)
    assert np.array_equal(np_(res["polygon_index"]), pp)
    assert np.array_equal(np_(res["point_index"]), pt)
    print(f"[full] subset hits={len(pp)}", flush=True)
    assert len(pp) > 1000
    # interleaved points (the benchmark's layout): identical tree, pairs and hits
    full = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, P)
    pxy = cs.Points.from_points_xy(torch.stack([w.x, w.y], dim=1).reshape(-1))
    pi2, tree2 = cs.quadtree_on_points(pxy, *bbox, w.scale, d, ms)
    assert torch.equal(i64(pi2), i64(pi))
    for k in ("key", "level", "is_internal_node", "length", "offset"):
        assert torch.equal(tree2[k].to(torch.int64) if k != "key" and k != "length" and
                           k != "offset" else i64(tree2[k]),
                           tree[k].to(torch.int64) if k != "key" and k != "length" and
                           k != "offset" else i64(tree[k])), k
    full2 = cs.quadtree_point_in_polygon(pairs, tree2, pi2, pxy, P)
    assert torch.equal(i64(full2["polygon_index"]), i64(full["polygon_index"]))
    assert torch.equal(i64(full2["point_index"]), i64(full["point_index"]))
    print(f"[full] interleaved layout identical, hits={len(full2)}", flush=True)


# ---------------------------------------------------------------------------
# boundary behaviour
# ---------------------------------------------------------------------------
def test_deterministic(cs):
    x, y, polys = scene(41, 300000)
    a = run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100 / 2**10, 10, 50, polys)
    b = run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100 / 2**10, 10, 50, polys)
    assert_same(a, b)


def test_non_default_stream(cs):
    import torch

    x, y, polys = scene(42, 100000)
    base = run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100 / 2**8, 8, 40, polys)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        other = run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100 / 2**8, 8, 40, polys)
    torch.cuda.synchronize()
    assert_same(other, base)


def test_memory_limit_raises_memory_error(cs):
    from cuspatial_amd import _capi

    x, y, _ = scene(43, 200000)
    L = _capi.lib()
    L.qpip_set_memory_limit(1 << 16)
    try:
        with pytest.raises(MemoryError):
            run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100 / 2**8, 8, 40)
    finally:
        L.qpip_set_memory_limit(0)
    run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100 / 2**8, 8, 40)


def test_default_pool_outputs_match_resource_outputs(cs):
    """mr == NULL: outputs come from the library pool and are released with qpip_free."""
    import torch

    from cuspatial_amd import _capi

    x, y, _ = scene(44, 50000)
    ref = run_gpu(cs, x, y, (0.0, 100.0, 0.0, 100.0), 100 / 2**7, 7, 30)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    xt = torch.as_tensor(x, device="cuda")
    yt = torch.as_tensor(y, device="cuda")
    outs = [ctypes.c_void_p() for _ in range(6)]
    m = ctypes.c_int64()
    L = _capi.lib()
    _capi.check(L.qpip_quadtree_on_points(
        ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(yt.data_ptr()), len(x), len(y), 0.0, 100.0,
        0.0, 100.0, 100 / 2**7, ctypes.c_int8(7), ctypes.c_int32(30), None, None,
        *[ctypes.byref(o) for o in outs], ctypes.byref(m)))
    sizes = [(len(x), np.uint32), (m.value, np.uint32), (m.value, np.uint8), (m.value, np.uint8),
             (m.value, np.uint32), (m.value, np.uint32)]
    got = []
    for o, (cnt, dt) in zip(outs, sizes):
        h = np.empty(cnt, dt)
        assert hip.hipMemcpy(h.ctypes.data, o.value, h.nbytes, 2) == 0
        got.append(h)
        L.qpip_free(o)
    for k, h in zip(["pi", "key", "level", "is_internal_node", "length", "offset"], got):
        assert np.array_equal(h.astype(np.int64), np.asarray(ref[k]).astype(np.int64)), k


def test_invalid_quadrant_index_is_an_error(cs):
    import torch

    x, y, polys = scene(45, 10000)
    pts = cs.Points(x, y)
    pi, tree = cs.quadtree_on_points(pts, 0.0, 100.0, 0.0, 100.0, 100 / 64, 6, 20)
    bad = cs.Table(bbox_offset=u32(torch.zeros(3, dtype=torch.int64, device="cuda")),
                   quad_offset=u32(torch.tensor([0, len(tree) + 5, 1], device="cuda")))
    with pytest.raises(RuntimeError, match="out of range"):
        cs.quadtree_point_in_polygon(bad, tree, pi, pts, cs.Polygons(*polys))


def test_float32_rejected(cs):
    import torch

    with pytest.raises(TypeError, match="float64"):
        cs.Points(torch.zeros(3, dtype=torch.float32), torch.zeros(3, dtype=torch.float32))


# ---------------------------------------------------------------------------
# brute-force point-in-polygon (every point x every polygon; pairwise)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", load_golden("pip_kats.json"), ids=lambda c: c["name"])
def test_point_in_polygon_bitmask_kats(cs, case):
    pts = np.array(case["points"], dtype=np.float64).reshape(-1, 2)
    v = np.array(case["vertices"], dtype=np.float64).reshape(-1, 2)
    P = cs.Polygons(case["part_offsets"], case["ring_offsets"], v[:, 0], v[:, 1], device="cuda")
    bm = cs.point_in_polygon_bitmask(cs.Points(pts[:, 0], pts[:, 1], device="cuda"), P)
    assert np_(bm).astype(np.int64).tolist() == case["expected"]


@pytest.mark.parametrize("case", load_golden("pairwise_pip_kats.json"), ids=lambda c: c["name"])
def test_pairwise_point_in_polygon_kats(cs, case):
    pts = np.array(case["points"], dtype=np.float64).reshape(-1, 2)
    v = np.array(case["vertices"], dtype=np.float64).reshape(-1, 2)
    P = cs.Polygons(case["part_offsets"], case["ring_offsets"], v[:, 0], v[:, 1], device="cuda")
    got = cs.pairwise_point_in_polygon(cs.Points(pts[:, 0], pts[:, 1], device="cuda"), P)
    assert np_(got).astype(np.int64).tolist() == case["expected"]


@pytest.mark.parametrize("seed,n,npoly", [(51, 1, 1), (52, 1000, 7), (53, 100003, 31),
                                          (54, 65, 31), (55, 300000, 12)])
@pytest.mark.parametrize("layout", ["columns", "xy"])
def test_point_in_polygon_bitmask_random(cs, oracle_mod, seed, n, npoly, layout):
    rng = np.random.default_rng(seed)
    part, ring, vx, vy = star_polygons(rng, npoly, 100.0, 5.0, 40.0, holes=0.3)
    x = rng.uniform(-5, 105, n)
    y = rng.uniform(-5, 105, n)
    # points on vertices and on edges' x ranges exercise the on-edge rule
    k = min(n, 50)
    x[:k] = vx[:k]
    y[:k] = vy[:k]
    pts = make_points(cs, x, y, layout)
    bm = cs.point_in_polygon_bitmask(pts, cs.Polygons(part, ring, vx, vy, device="cuda"))
    ref = oracle_mod.point_in_polygon_bitmask(x, y, part, ring, vx, vy)
    assert np.array_equal(np_(bm).astype(np.int64), ref.astype(np.int64))
    tbl = cs.point_in_polygon(pts, cs.Polygons(part, ring, vx, vy, device="cuda"))
    assert tbl.columns == list(range(npoly))
    for j in range(npoly):
        assert np.array_equal(np_(tbl[j]), ((ref >> j) & 1).astype(bool))


@pytest.mark.parametrize("seed,n", [(61, 1), (62, 5000), (63, 200001)])
def test_pairwise_point_in_polygon_random(cs, oracle_mod, seed, n):
    rng = np.random.default_rng(seed)
    part, ring, vx, vy = star_polygons(rng, n, 1000.0, 5.0, 40.0, holes=0.3)
    cx = np.array([vx[ring[part[i]]:ring[part[i] + 1]].mean() for i in range(min(n, 2000))])
    x = rng.uniform(0, 1000, n)
    y = rng.uniform(0, 1000, n)
    m = len(cx)
    x[:m] = cx + rng.normal(0, 5, m)  # many points near their own polygon
    y[:m] = np.array([vy[ring[part[i]]:ring[part[i] + 1]].mean() for i in range(m)])
    got = cs.pairwise_point_in_polygon(cs.Points(x, y, device="cuda"),
                                       cs.Polygons(part, ring, vx, vy, device="cuda"))
    ref = oracle_mod.pairwise_point_in_polygon(x, y, part, ring, vx, vy)
    assert np.array_equal(np_(got).astype(np.uint8), ref)
    assert int(ref.sum()) > 0 or n < 10


def test_point_in_polygon_errors(cs):
    rng = np.random.default_rng(71)
    part, ring, vx, vy = star_polygons(rng, 32, 100.0, 5.0, 20.0)
    pts = cs.Points(np.zeros(3), np.zeros(3), device="cuda")
    with pytest.raises(RuntimeError, match="Number of polygons cannot exceed 31"):
        cs.point_in_polygon_bitmask(pts, cs.Polygons(part, ring, vx, vy, device="cuda"))
    with pytest.raises(RuntimeError, match="Must pass in the same number of points as polygons."):
        cs.pairwise_point_in_polygon(pts, cs.Polygons(part, ring, vx, vy, device="cuda"))
