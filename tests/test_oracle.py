"""Pins the oracle (CPU restatement of the reference path) before it is trusted.

Each function of the oracle is checked against
  * the reference's own golden vectors (tests/golden, extracted from the
    reference's gtests), and
  * independent formulations that a plausible transliteration mistake would
    fail: bit-loop Morton codes, ULP arithmetic built from numpy.nextafter, a
    top-down-from-definition quadtree, a brute-force overlap join, brute-force
    all-pairs point-in-polygon, numpy min/max boxes, and a ray-casting
    reference for points away from polygon edges.
"""

import numpy as np
import pytest

from conftest import load_golden, star_polygons


# ---------------------------------------------------------------------------
# golden vectors
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", load_golden("quadtree_cases.json"), ids=lambda c: c["name"])
def test_quadtree_golden(oracle_mod, case):
    p = np.array(case["points"], dtype=np.float64).reshape(-1, 2)
    v1, v2 = case["vertex_1"], case["vertex_2"]
    pi, key, level, internal, length, offset = oracle_mod.quadtree_on_points(
        p[:, 0], p[:, 1], v1[0], v2[0], v1[1], v2[1], case["scale"], case["max_depth"],
        case["max_size"])
    assert len(pi) == len(p)
    assert key.tolist() == case["key"]
    assert level.tolist() == case["level"]
    assert internal.tolist() == case["is_internal_node"]
    assert length.tolist() == case["length"]
    assert offset.tolist() == case["offset"]


def test_join_and_pip_golden(oracle_mod):
    g = load_golden("join_pip_small.json")
    p = np.array(g["points"])
    v = np.array(g["vertices"])
    tree = oracle_mod.quadtree_on_points(p[:, 0], p[:, 1], 0, 8, 0, 8, g["scale"], g["max_depth"],
                                         g["max_size"])
    pi, t5 = tree[0], tree[1:]
    bb = oracle_mod.polygon_bounding_boxes(g["part_offsets"], g["ring_offsets"], v[:, 0], v[:, 1])
    jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(t5, bb, 0, 8, 0, 8, g["scale"],
                                                         g["max_depth"])
    assert jb.tolist() == g["expected_join_bbox_offset"]
    assert jq.tolist() == g["expected_join_quad_offset"]
    pp, pt = oracle_mod.quadtree_point_in_polygon((jb, jq), t5, pi, p[:, 0], p[:, 1],
                                                  g["part_offsets"], g["ring_offsets"], v[:, 0],
                                                  v[:, 1])
    assert pp.tolist() == g["expected_polygon_index"]
    assert pt.tolist() == g["expected_point_index"]


@pytest.mark.parametrize("case", load_golden("pairwise_pip_kats.json"), ids=lambda c: c["name"])
def test_pairwise_pip_kats(oracle_mod, case):
    """The reference's pairwise vectors (pairwise_point_in_polygon_test.cu)."""
    pts = np.array(case["points"], dtype=np.float64).reshape(-1, 2)
    v = np.array(case["vertices"], dtype=np.float64).reshape(-1, 2)
    got = oracle_mod.pairwise_point_in_polygon(pts[:, 0], pts[:, 1], case["part_offsets"],
                                               case["ring_offsets"], v[:, 0], v[:, 1])
    assert got.tolist() == case["expected"]


def test_pairwise_matches_bitmask_diagonal(oracle_mod):
    """Pairwise result i equals bit i of the brute-force bitmask of point i."""
    rng = np.random.default_rng(5)
    from conftest import star_polygons

    part, ring, vx, vy = star_polygons(rng, 20, 100.0, 5.0, 30.0, holes=0.3)
    n = len(part) - 1
    x = rng.uniform(0, 100, n)
    y = rng.uniform(0, 100, n)
    bm = oracle_mod.point_in_polygon_bitmask(x, y, part, ring, vx, vy)
    pw = oracle_mod.pairwise_point_in_polygon(x, y, part, ring, vx, vy)
    assert pw.tolist() == [(int(b) >> i) & 1 for i, b in enumerate(bm)]


def _polys(polygons):
    part, ring, verts = [0], [0], []
    for r in polygons:
        verts += r
        ring.append(len(verts))
        part.append(len(ring) - 1)
    v = np.array(verts, dtype=np.float64).reshape(-1, 2)
    return part, ring, v


@pytest.mark.parametrize("case", [c for c in load_golden("contains_properly_kats.json")
                                  if len(c["points"]) == 1 and len(c["polygons"]) == 1],
                         ids=lambda c: "%s_%s" % (c["source"].split(":")[-1], c["points"][0]))
def test_contains_properly_kats(oracle_mod, case):
    """The reference's contains_properly expectations for one point against one
    polygon (tests/binpreds/test_contains_properly.py): the brute-force test the
    caller uses for up to 31 polygons is the predicate itself."""
    part, ring, v = _polys(case["polygons"])
    pts = np.array(case["points"], dtype=np.float64)
    got = oracle_mod.point_in_polygon_bitmask(pts[:, 0], pts[:, 1], part, ring, v[:, 0], v[:, 1])
    assert [bool(b & 1) for b in got] == case["expected"]


@pytest.mark.parametrize("case", load_golden("pip_kats.json"), ids=lambda c: c["name"])
def test_pip_kats(oracle_mod, case):
    pts = np.array(case["points"]).reshape(-1, 2)
    v = np.array(case["vertices"])
    bm = oracle_mod.point_in_polygon_bitmask(pts[:, 0], pts[:, 1], case["part_offsets"],
                                             case["ring_offsets"], v[:, 0], v[:, 1])
    assert bm.tolist() == case["expected"]


# ---------------------------------------------------------------------------
# Morton codes: bit loop, exhaustive over one coordinate, random pairs
# ---------------------------------------------------------------------------
def _spread(v):
    v = np.asarray(v, dtype=np.uint64)
    out = np.zeros_like(v)
    for i in range(16):
        out |= ((v >> np.uint64(i)) & np.uint64(1)) << np.uint64(2 * i)
    return out


def test_z_order_matches_bit_loop(oracle_mod):
    vals = np.arange(65536, dtype=np.uint64)
    want = _spread(vals)
    got = np.array([oracle_mod.z_order(int(v), 0) for v in range(0, 65536, 1)], dtype=np.uint64)
    assert np.array_equal(got, want)
    rng = np.random.default_rng(3)
    for x, y in rng.integers(0, 65536, (2000, 2)):
        k = oracle_mod.z_order(int(x), int(y))
        assert k == int((_spread(y) << np.uint64(1)) | _spread(x))
        assert oracle_mod.z_order_x(k) == x and oracle_mod.z_order_y(k) == y


# ---------------------------------------------------------------------------
# float_equal: ULP distance built from nextafter
# ---------------------------------------------------------------------------
def _step(a, n):
    for _ in range(abs(n)):
        a = np.nextafter(a, np.inf if n > 0 else -np.inf)
    return float(a)


@pytest.mark.parametrize("base", [1.0, -3.5, 1e300, 1e-300, 5e-324, -5e-324, 123456.789])
def test_float_equal_ulps(oracle_mod, base):
    for n in range(-6, 7):
        assert oracle_mod.float_equal(base, _step(base, n)) == (abs(n) <= 4), n


def test_float_equal_special(oracle_mod):
    fe = oracle_mod.float_equal
    assert fe(0.0, -0.0)
    assert fe(_step(0.0, 2), _step(0.0, -2))        # 4 ULP across zero
    assert not fe(_step(0.0, 3), _step(0.0, -2))    # 5 ULP across zero
    assert not fe(np.nan, np.nan) and not fe(np.nan, 1.0)
    assert fe(np.inf, np.inf)
    assert fe(np.inf, np.finfo(np.float64).max)     # adjacent in the biased order
    assert not fe(1.0, -1.0)


# ---------------------------------------------------------------------------
# point-in-polygon against ray casting for points away from every edge,
# and on-edge special cases
# ---------------------------------------------------------------------------
def _raycast(px, py, rings):
    inside = False
    for vx, vy in rings:
        n = len(vx)
        for i in range(n):
            x1, y1, x2, y2 = vx[i - 1], vy[i - 1], vx[i], vy[i]
            if (y1 > py) != (y2 > py):
                xint = x1 + (py - y1) * (x2 - x1) / (y2 - y1)
                if px < xint:
                    inside = not inside
    return inside


def _edge_distance(px, py, rings):
    d = np.inf
    for vx, vy in rings:
        for i in range(len(vx)):
            x1, y1, x2, y2 = vx[i - 1], vy[i - 1], vx[i], vy[i]
            dx, dy = x2 - x1, y2 - y1
            L = dx * dx + dy * dy
            t = 0.0 if L == 0 else min(1.0, max(0.0, ((px - x1) * dx + (py - y1) * dy) / L))
            d = min(d, np.hypot(px - (x1 + t * dx), py - (y1 + t * dy)))
    return d


def test_pip_matches_ray_casting(oracle_mod):
    rng = np.random.default_rng(11)
    part, ring, vx, vy = star_polygons(rng, 40, 100.0, 5.0, 20.0, 3, 16, holes=0.5)
    pts = rng.uniform(0, 100, (3000, 2))
    for j in range(len(part) - 1):
        rings = [(vx[ring[r]:ring[r + 1]], vy[ring[r]:ring[r + 1]])
                 for r in range(part[j], part[j + 1])]
        for px, py in pts[:150]:
            if _edge_distance(px, py, rings) < 1e-6:
                continue
            got = oracle_mod.is_point_in_polygon(px, py, ring, part[j], part[j + 1], vx, vy)
            assert got == _raycast(px, py, rings)


def test_pip_vertices_and_vertical_extension_are_outside(oracle_mod):
    # every vertex lies on an edge -> outside
    ring = np.array([0, 5], np.uint32)
    vx = np.array([0.0, 4.0, 4.0, 0.0, 0.0])
    vy = np.array([0.0, 0.0, 3.0, 3.0, 0.0])
    for x, y in zip(vx, vy):
        assert not oracle_mod.is_point_in_polygon(x, y, ring, 0, 1, vx, vy)
    # is_point_in_polygon.cuh:56-64 checks collinearity against the infinite line
    # and only the x range of the edge, so a point on the vertical line through a
    # vertical edge counts as "on edge" even when it is off the segment. In the
    # L-shape below, (2, 1) is geometrically inside but shares x with the
    # vertical edge (2,2)-(2,4): the reference reports it outside.
    ring = np.array([0, 7], np.uint32)
    lx = np.array([0.0, 4.0, 4.0, 2.0, 2.0, 0.0, 0.0])
    ly = np.array([0.0, 0.0, 2.0, 2.0, 4.0, 4.0, 0.0])
    assert not oracle_mod.is_point_in_polygon(2.0, 1.0, ring, 0, 1, lx, ly)
    assert oracle_mod.is_point_in_polygon(1.0, 1.0, ring, 0, 1, lx, ly)
    assert oracle_mod.is_point_in_polygon(3.0, 1.0, ring, 0, 1, lx, ly)
    assert oracle_mod.is_point_in_polygon(1.0, 3.0, ring, 0, 1, lx, ly)
    assert not oracle_mod.is_point_in_polygon(3.0, 3.0, ring, 0, 1, lx, ly)
    # The same holds across rings: the on-edge hit of any ring makes the whole
    # polygon false (is_point_in_polygon.cuh:93-96). A point inside the outer
    # ring, on the vertical line through a hole's edge but above the hole, is
    # outside; a point beside that line is inside.
    hr = np.array([0, 5, 10], np.uint32)
    hx = np.array([0.0, 10.0, 10.0, 0.0, 0.0, 2.0, 4.0, 4.0, 2.0, 2.0])
    hy = np.array([0.0, 0.0, 10.0, 10.0, 0.0, 2.0, 2.0, 4.0, 4.0, 2.0])
    for x, y in ((2.0, 7.0), (4.0, 7.0), (2.0, 1.0), (4.0, 0.5), (3.0, 3.0)):
        assert not oracle_mod.is_point_in_polygon(x, y, hr, 0, 2, hx, hy), (x, y)
    for x, y in ((3.0, 7.0), (1.0, 3.0), (2.5, 9.0), (np.nextafter(2.0, 3.0), 7.0)):
        assert oracle_mod.is_point_in_polygon(x, y, hr, 0, 2, hx, hy), (x, y)


# ---------------------------------------------------------------------------
# quadtree against a top-down construction from the definition
# ---------------------------------------------------------------------------
def naive_quadtree(x, y, xmin, xmax, ymin, ymax, scale, d, max_size):
    """Nodes are the key prefixes at levels 0..d-1 whose ancestors all hold more
    than max_size points; internal iff more than max_size points and level <= d-2;
    level-major, key order; internal offset = position of first child, leaf
    offset = first position in the stable key order."""
    xmin, xmax = min(xmin, xmax), max(xmin, xmax)
    ymin, ymax = min(ymin, ymax), max(ymin, ymax)
    max_size = max(1, max_size)
    d = max(0, min(15, d))
    scale = max(scale, max(xmax - xmin, ymax - ymin) / ((1 << d) + 2))
    out = (x < xmin) | (x > xmax) | (y < ymin) | (y > ymax)
    cx = np.where(out, 0, np.trunc((x - xmin) / scale)).astype(np.uint64)
    cy = np.where(out, 0, np.trunc((y - ymin) / scale)).astype(np.uint64)
    keys = (_spread(cy) << np.uint64(1)) | _spread(cx)
    keys = np.where(out, np.uint64((1 << (2 * d)) - 1), keys)
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    nodes = []  # (key, level, start, count)
    top_shift = 2 * (d - 1) if d >= 1 else 0
    lvl0 = sk >> np.uint64(top_shift)
    uniq, starts, counts = np.unique(lvl0, return_index=True, return_counts=True)
    frontier = [(int(k), 0, int(s), int(c)) for k, s, c in zip(uniq, starts, counts)]
    levels = [frontier]
    for lvl in range(0, d - 1):
        nxt = []
        for k, l, s, c in levels[-1]:
            if c > max_size:
                sub = sk[s:s + c] >> np.uint64(2 * (d - 2 - lvl))
                u, st, ct = np.unique(sub, return_index=True, return_counts=True)
                nxt += [(int(uu), lvl + 1, s + int(ss), int(cc)) for uu, ss, cc in zip(u, st, ct)]
        levels.append(nxt)
    flat = [n for lv in levels for n in lv]
    key, level, internal, length, offset = [], [], [], [], []
    pos = {}
    for i, n in enumerate(flat):
        pos[(n[1], n[0])] = i
    for k, l, s, c in flat:
        is_int = c > max_size and l <= d - 2
        key.append(k)
        level.append(l)
        internal.append(is_int)
        if is_int:
            first = pos[(l + 1, min(kk for kk, ll, ss, cc in levels[l + 1] if kk >> 2 == k))]
            length.append(sum(1 for kk, ll, ss, cc in levels[l + 1] if kk >> 2 == k))
            offset.append(first)
        else:
            length.append(c)
            offset.append(s)
    return order.astype(np.uint32), key, level, internal, length, offset


@pytest.mark.parametrize("seed,n,d,max_size", [
    (0, 50, 1, 1), (1, 200, 2, 3), (2, 500, 3, 12), (3, 2000, 5, 20), (4, 3000, 7, 125),
    (5, 5000, 8, 7), (6, 4000, 15, 64), (7, 100, 0, 1), (8, 3000, 4, 1), (9, 1000, 6, 1000)])
def test_quadtree_matches_definition(oracle_mod, seed, n, d, max_size):
    rng = np.random.default_rng(seed)
    # clustered points, some outside the bbox, some duplicates
    x = np.concatenate([rng.normal(30, 8, n // 2), rng.uniform(-5, 100, n - n // 2)])
    y = np.concatenate([rng.normal(60, 5, n // 2), rng.uniform(0, 105, n - n // 2)])
    x[:10] = x[10:20]
    y[:10] = y[10:20]
    scale = 100.0 / (1 << d)
    got = oracle_mod.quadtree_on_points(x, y, 0, 100, 0, 100, scale, d, max_size)
    want = naive_quadtree(x, y, 0, 100, 0, 100, scale, d, max_size)
    assert np.array_equal(got[0], want[0])
    for g, w in zip(got[1:], want[1:]):
        assert np.asarray(g).tolist() == [int(v) if not isinstance(v, bool) else v for v in w]


# ---------------------------------------------------------------------------
# bounding boxes, join and PIP against brute force
# ---------------------------------------------------------------------------
def test_bbox_bruteforce(oracle_mod):
    rng = np.random.default_rng(5)
    part, ring, vx, vy = star_polygons(rng, 60, 50.0, 1.0, 5.0, holes=0.3)
    got = oracle_mod.polygon_bounding_boxes(part, ring, vx, vy)
    for j in range(len(part) - 1):
        a, b = ring[part[j]], ring[part[j + 1]]
        assert got[0][j] == vx[a:b].min() and got[2][j] == vx[a:b].max()
        assert got[1][j] == vy[a:b].min() and got[3][j] == vy[a:b].max()


def _undilate(k):
    out = 0
    for i in range(16):
        out |= ((k >> (2 * i)) & 1) << i
    return out


@pytest.mark.parametrize("seed,d,max_size", [(0, 3, 4), (1, 5, 10), (2, 7, 30)])
def test_join_bruteforce(oracle_mod, seed, d, max_size):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 64, 3000)
    y = rng.uniform(0, 64, 3000)
    scale = 64.0 / (1 << d)  # power of two: products are exact
    pi, key, level, internal, length, offset = oracle_mod.quadtree_on_points(
        x, y, 0, 64, 0, 64, scale, d, max_size)
    part, ring, vx, vy = star_polygons(rng, 25, 64.0, 0.5, 6.0)
    bb = oracle_mod.polygon_bounding_boxes(part, ring, vx, vy)
    jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(
        (key, level, internal, length, offset), bb, 0, 64, 0, 64, scale, d)
    parent = {}
    for i in range(len(key)):
        if internal[i]:
            for c in range(offset[i], offset[i] + length[i]):
                parent[c] = i

    def box(i):
        ls = scale * (1 << (d - 1 - int(level[i])))
        kx, ky = _undilate(int(key[i])), _undilate(int(key[i]) >> 1)
        return kx * ls, ky * ls, (kx + 1) * ls, (ky + 1) * ls

    def overlap(i, b):
        x0, y0, x1, y1 = box(i)
        return not (x0 > bb[2][b] or x1 < bb[0][b] or y0 > bb[3][b] or y1 < bb[1][b])

    want = []
    for i in range(len(key)):
        if internal[i]:
            continue
        for b in range(len(bb[0])):
            j, ok = i, True
            while ok and j is not None:
                ok = overlap(j, b)
                j = parent.get(j)
            if ok:
                want.append((int(offset[i]), b, i))
    want.sort()
    assert list(zip(jb.tolist(), jq.tolist())) == [(b, i) for _, b, i in want]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_quadtree_pip_equals_bruteforce(oracle_mod, seed):
    """The reference's large-test method (quadtree_point_in_polygon_test_large.cu:170-224)."""
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.normal(50, 12, 2000), rng.uniform(0, 100, 2000)]).clip(0, 99.99)
    y = np.concatenate([rng.normal(40, 9, 2000), rng.uniform(0, 100, 2000)]).clip(0, 99.99)
    part, ring, vx, vy = star_polygons(rng, 30, 100.0, 3.0, 15.0, holes=0.3)
    ok = (vx >= 0).all() and (vx <= 100).all()
    d, max_size = 6, 40
    scale = 100.0 / (1 << d)
    tree = oracle_mod.quadtree_on_points(x, y, 0, 100, 0, 100, scale, d, max_size)
    bb = oracle_mod.polygon_bounding_boxes(part, ring, vx, vy)
    jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(tree[1:], bb, 0, 100, 0, 100, scale, d)
    pp, pt = oracle_mod.quadtree_point_in_polygon((jb, jq), tree[1:], tree[0], x, y, part, ring,
                                                  vx, vy)
    got = sorted(zip(tree[0][pt].tolist(), pp.tolist()))
    bp, bpt = oracle_mod.all_pairs_bruteforce(x, y, part, ring, vx, vy)
    want = sorted(zip(bpt.tolist(), bp.tolist()))
    if ok:
        assert got == want
    else:  # polygons poking out of the bbox may miss points outside it
        assert set(got) <= set(want)
    assert len(got) > 50


def _ulps(a, b):
    a = np.asarray(a, np.float64).view(np.int64)
    b = np.asarray(b, np.float64).view(np.int64)
    return np.abs(a - b)


def test_nearest_linestring_golden(oracle_mod):
    """quadtree_point_to_nearest_linestring_test_small.cu: the join over expanded
    linestring boxes and, per point, the nearest linestring and the distance,
    both exactly (the reference test allows 4 ULP; the fma-contracted dot
    products reproduce its values bit for bit)."""
    g = load_golden("nearest_linestring_small.json")
    p = np.array(g["points"], dtype=np.float64)
    v = np.array(g["vertices"], dtype=np.float64)
    x0, x1, y0, y1 = g["bbox"]
    pi, key, level, internal, length, offset = oracle_mod.quadtree_on_points(
        p[:, 0], p[:, 1], x0, x1, y0, y1, g["scale"], g["max_depth"], g["max_size"])
    tree = (key, level, internal, length, offset)
    bb = oracle_mod.linestring_bounding_boxes(g["linestring_offsets"], v[:, 0], v[:, 1],
                                              g["expansion_radius"])
    jl, jq = oracle_mod.join_quadtree_and_bounding_boxes(tree, bb, x0, x1, y0, y1, g["scale"],
                                                         g["max_depth"])
    assert jl.tolist() == g["expected_join_linestring_index"]
    assert jq.tolist() == g["expected_join_quad_index"]
    pt, ls, d = oracle_mod.quadtree_point_to_nearest_linestring(
        (jl, jq), tree, pi, p[:, 0], p[:, 1], g["linestring_offsets"], v[:, 0], v[:, 1])
    assert pt.tolist() == g["expected_point_index"]
    assert ls.tolist() == g["expected_linestring_index"]
    assert _ulps(d, g["expected_distance"]).max() == 0


def test_nearest_linestring_bruteforce(oracle_mod):
    """Random linestrings and points: each point's row is the nearest of the
    linestrings paired with its leaf (brute force over those candidates, with the
    reference's zero-distance and tie rules), by an independent numpy distance."""
    rng = np.random.default_rng(17)
    n, L = 3000, 40
    x, y = rng.uniform(0, 100, n), rng.uniform(0, 100, n)
    offs, lx, ly = [0], [], []
    for _ in range(L):
        k = int(rng.integers(2, 7))
        cx, cy = rng.uniform(0, 100, 2)
        lx += list(cx + rng.normal(0, 6, k))
        ly += list(cy + rng.normal(0, 6, k))
        offs.append(len(lx))
    lx, ly = np.array(lx), np.array(ly)
    pi, key, level, internal, length, offset = oracle_mod.quadtree_on_points(
        x, y, 0, 100, 0, 100, 100 / 32, 5, 40)
    tree = (key, level, internal, length, offset)
    bb = oracle_mod.linestring_bounding_boxes(offs, lx, ly, 8.0)
    jl, jq = oracle_mod.join_quadtree_and_bounding_boxes(tree, bb, 0, 100, 0, 100, 100 / 32, 5)
    pt, ls, d = oracle_mod.quadtree_point_to_nearest_linestring(
        (jl, jq), tree, pi, x, y, offs, lx, ly)
    assert pt.tolist() == list(range(n))

    def dist(px, py, j):
        best = np.inf
        for i in range(offs[j], offs[j + 1] - 1):
            ax, ay, bx, by = lx[i], ly[i], lx[i + 1], ly[i + 1]
            t = ((px - ax) * (bx - ax) + (py - ay) * (by - ay)) / ((bx - ax) ** 2 + (by - ay) ** 2)
            t = min(max(t, 0.0), 1.0)
            best = min(best, (px - ax - t * (bx - ax)) ** 2 + (py - ay - t * (by - ay)) ** 2)
        return np.sqrt(best)

    cand = {}
    for l_, q in zip(jl, jq):
        cand.setdefault(int(q), []).append(int(l_))
    leaf_of = np.full(n, -1)
    for q in cand:
        leaf_of[offset[q]:offset[q] + length[q]] = q
    covered = 0
    for pos in range(n):
        q = leaf_of[pos]
        if q < 0:
            assert ls[pos] == 0xFFFFFFFF and d[pos] == 0.0
            continue
        covered += 1
        idx = pi[pos]
        ds = [(dist(x[idx], y[idx], j), j) for j in cand[q]]
        nz = [t for t in ds if t[0] != 0]
        want = min(nz) if nz else ds[-1]
        assert ls[pos] == want[1], pos
        assert abs(d[pos] - want[0]) <= 1e-9 * max(1.0, want[0])
    assert covered > n // 2
