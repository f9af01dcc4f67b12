"""The quadtree point-in-polygon test at the predicate's boundaries.

Points a few ULPs above, below and beside polygon vertices; on the lines of
horizontal edges beyond their ends; on the vertical lines through vertical edges
far above and below them (the reference's on-edge test holds there for any y:
run == 0 makes both cross products zero); tight clusters hugging a vertex's y
from one side; thin slivers with nearly horizontal and nearly vertical edges.
These are where the collinearity filter and the crossing rule decide, and where
any shortcut that passes over edges would go wrong.

Bar: the hits equal the oracle's bit for bit, and the per-point polygon sets
equal the brute-force bitmask's (GPU and oracle).
"""

import numpy as np
import pytest

from test_gpu_parity import assert_same, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cs():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd

    return cuspatial_amd


def ulps(v, k):
    """v moved k ULPs (k may be negative)."""
    out = np.float64(v)
    step = np.inf if k > 0 else -np.inf
    for _ in range(abs(k)):
        out = np.nextafter(out, step)
    return out


def scene(seed, ext=64.0):
    rng = np.random.default_rng(seed)
    part, ring, vx, vy = [0], [0], [], []

    def add_ring(xs, ys):
        xs, ys = list(xs) + [xs[0]], list(ys) + [ys[0]]
        vx.extend(xs)
        vy.extend(ys)
        ring.append(ring[-1] + len(xs))
        part.append(part[-1] + 1)

    # axis-aligned rectangles: vertical and horizontal edges
    for _ in range(12):
        x0, y0 = rng.uniform(2, ext - 12, 2)
        w, h = rng.uniform(0.5, 8, 2)
        add_ring([x0, x0 + w, x0 + w, x0], [y0, y0, y0 + h, y0 + h])
    # stars and thin slivers (nearly horizontal / nearly vertical edges)
    for _ in range(12):
        cx, cy = rng.uniform(8, ext - 8, 2)
        k = int(rng.integers(3, 14))
        a = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = rng.uniform(1, 6) * rng.uniform(0.6, 1.4, k)
        add_ring(cx + r * np.cos(a), cy + r * np.sin(a))
    for _ in range(6):
        x0, y0 = rng.uniform(4, ext - 20, 2)
        add_ring([x0, x0 + 15, x0 + 7.5], [y0, ulps(y0, 3), y0 + 1e-9])
        add_ring([x0, ulps(x0, 2), x0 + 1e-9], [y0, y0 + 15, y0 + 7.5])
    vx, vy = np.array(vx), np.array(vy)
    # points: around every vertex (a few ULPs in y and x), on vertical lines above
    # and below, on horizontal lines beyond the ends, tight clusters, background
    xs, ys = [], []
    for px_, py_ in zip(vx, vy):
        for k in (-8, -4, -2, -1, 0, 1, 2, 4, 8):
            xs.append(px_)
            ys.append(ulps(py_, k))
        for k in (-2, 2):
            xs.append(ulps(px_, k))
            ys.append(py_)
        for dy in (-20.0, -3.0, 3.0, 20.0):  # the vertical line through the vertex
            xs.append(px_)
            ys.append(py_ + dy)
        for dx in (-5.0, 5.0):  # the horizontal line through the vertex
            xs.append(px_ + dx)
            ys.append(py_)
    for _ in range(40):  # clusters hugging a vertex's y from one side
        i = int(rng.integers(len(vx)))
        side = 1 if rng.uniform() < 0.5 else -1
        m = 64
        xs.extend(vx[i] + rng.uniform(-0.5, 0.5, m))
        ys.extend([ulps(vy[i], side * int(rng.integers(1, 6))) for _ in range(m)])
    xs.extend(rng.uniform(0, ext, 3000))
    ys.extend(rng.uniform(0, ext, 3000))
    x = np.clip(np.array(xs), 0.0, np.nextafter(ext, 0))
    y = np.clip(np.array(ys), 0.0, np.nextafter(ext, 0))
    return x, y, (np.array(part, np.int32), np.array(ring, np.int32), vx, vy)


@pytest.mark.parametrize("seed,d,max_size", [(1, 6, 16), (2, 8, 64), (3, 4, 256), (4, 10, 4),
                                              (5, 7, 1)])
def test_boundaries_match_oracle(cs, oracle_mod, seed, d, max_size):
    x, y, polys = scene(seed)
    bbox = (0.0, 64.0, 0.0, 64.0)
    scale = 64.0 / (1 << d)
    g = run_gpu(cs, x, y, bbox, scale, d, max_size, polys)
    o = run_oracle(oracle_mod, x, y, bbox, scale, d, max_size, polys)
    assert_same(g, o)
    assert len(o["polygon_index"]) > 0


def test_boundaries_match_bitmask(cs, oracle_mod):
    """Per point, the polygons the quadtree path finds equal the brute-force
    bitmask's (31 polygons), on the GPU and in the oracle."""
    x, y, (part, ring, vx, vy) = scene(9)
    part = part[:32]
    ring = ring[:int(part[-1]) + 1]
    vx, vy = vx[:int(ring[-1])], vy[:int(ring[-1])]
    polys = (part, ring, vx, vy)
    g = run_gpu(cs, x, y, (0.0, 64.0, 0.0, 64.0), 64.0 / 256, 8, 32, polys)
    bits = np.zeros(len(x), np.int64)
    for poly, pos in zip(g["polygon_index"], g["point_index"]):
        bits[g["pi"][pos]] |= 1 << int(poly)
    P = cs.Polygons(part, ring, vx, vy, device="cuda")
    bm = cs.point_in_polygon_bitmask(cs.Points(x, y, device="cuda"), P)
    assert np.array_equal(bits, bm.cpu().numpy().astype(np.int64))
    ob = oracle_mod.point_in_polygon_bitmask(x, y, part, ring, vx, vy)
    assert np.array_equal(bits, np.asarray(ob).astype(np.int64))


@pytest.mark.parametrize("waves,stage,min_pairs,xcd_run,box,group", [
    (2, 64, 2, 16, 1, 0), (8, 64, 2, 1, 0, 0), (4, 128, 2, 4, 1, 0), (8, 128, 1, 64, 1, 0),
    (2, 128, 16, 1, 0, 0), (2, 128, 1, 256, 1, 0), (2, 128, 1, 16, 1, 1), (2, 64, 1, 16, 1, 3),
    (4, 128, 1, 16, 0, 16), (8, 64, 1, 16, 1, 4096)])
def test_tuning_does_not_change_results(cs, oracle_mod, waves, stage, min_pairs, xcd_run, box,
                                        group):
    """The PIP test's workgroup size, staging depth, segment length, XCD run, box
    filter and polygon-major grouping change only the time: every setting gives
    the oracle's pairs on the boundary scene."""
    from cuspatial_amd import autotune

    x, y, polys = scene(21)
    bbox = (0.0, 64.0, 0.0, 64.0)
    saved = autotune.current()
    try:
        autotune.apply({"pip_waves_per_block": waves, "pip_stage_edges": stage,
                        "pip_min_pairs_per_wave": min_pairs, "pip_xcd_run": xcd_run,
                        "pip_box_filter": box, "pip_polygon_group": group})
        g = run_gpu(cs, x, y, bbox, 64.0 / 64, 6, 8, polys)
    finally:
        autotune.apply(saved)
    o = run_oracle(oracle_mod, x, y, bbox, 64.0 / 64, 6, 8, polys)
    assert_same(g, o)


def test_hole_vertical_lines_match_oracle(cs, oracle_mod):
    """Squares with square holes: points inside an outer ring on the vertical lines
    through a hole's edges (above and below the hole) are "on edge" for the hole,
    so the reference reports the whole polygon false there
    (is_point_in_polygon.cuh:56-65, :93-96); one ULP beside the line they are
    inside. Any per-ring shortcut would get these wrong. Bar: the oracle's pairs
    bit for bit, and the hand-picked answers."""
    part, ring, vx, vy = [0], [0], [], []
    xs, ys, want_in, want_out = [], [], [], []
    rng = np.random.default_rng(11)
    for k in range(16):
        ox, oy = 4.0 * (k % 4) + 1.0, 4.0 * (k // 4) + 1.0  # outer [ox, ox+3]^2
        hx0, hy0 = ox + 1.0, oy + 1.0                      # hole [hx0, hx0+1]^2
        for rx, ry in (([ox, ox + 3, ox + 3, ox, ox], [oy, oy, oy + 3, oy + 3, oy]),
                       ([hx0, hx0 + 1, hx0 + 1, hx0, hx0], [hy0, hy0, hy0 + 1, hy0 + 1, hy0])):
            vx.extend(rx)
            vy.extend(ry)
            ring.append(ring[-1] + 5)
        part.append(part[-1] + 2)
        for hx in (hx0, hx0 + 1.0):
            for py_ in (oy + 0.5, oy + 2.5):  # below / above the hole, inside the outer ring
                want_out.append(len(xs))
                xs.append(hx)
                ys.append(py_)
                want_in.append(len(xs))
                xs.append(np.nextafter(hx, ox) if hx == hx0 else np.nextafter(hx, ox + 3.0))
                ys.append(py_)
        xs.extend(rng.uniform(ox, ox + 3.0, 40))
        ys.extend(rng.uniform(oy, oy + 3.0, 40))
    x, y = np.array(xs), np.array(ys)
    polys = (np.array(part, np.int32), np.array(ring, np.int32), np.array(vx), np.array(vy))
    bbox = (0.0, 17.0, 0.0, 17.0)
    for d, max_size in ((5, 8), (3, 64)):
        scale = 17.0 / (1 << d)
        g = run_gpu(cs, x, y, bbox, scale, d, max_size, polys)
        o = run_oracle(oracle_mod, x, y, bbox, scale, d, max_size, polys)
        assert_same(g, o)
        hit = np.zeros(len(x), bool)
        hit[g["pi"][g["point_index"]]] = True
        assert not hit[want_out].any()
        assert hit[want_in].all()


@pytest.mark.parametrize("k", [-20, 0, 30])
def test_box_filter_margins_match_oracle(cs, oracle_mod, k):
    """Points around each polygon's vertex box (a few ULPs out, around the box
    filter's margin and beyond, on all four sides), including points left of the
    box that the reference's rounding calls inside (test_pip_filter.py): the hits
    equal the oracle's bit for bit, and those points are among them."""
    from test_pip_filter import _probe_points, _shapes, ring_box

    rng = np.random.default_rng(40 + k)
    s = 2.0 ** k
    shapes = _shapes(rng, s)
    part, ring, vx, vy, xs, ys = [0], [0], [], [], [], []
    for sx, sy in shapes:
        vx.extend(sx)
        vy.extend(sy)
        ring.append(ring[-1] + len(sx))
        part.append(part[-1] + 1)
        px, py = _probe_points(ring_box(sx, sy), sx, sy, rng, n_rand=40)
        xs.extend(px)
        ys.extend(py)
    quirk = (np.nextafter(0.5 * s, 0.0), np.nextafter(0.5 * s, 0.0))
    xs.append(quirk[0])
    ys.append(quirk[1])
    xs.extend(rng.uniform(-2 * s, 2 * s, 2000))
    ys.extend(rng.uniform(-2 * s, 2 * s, 2000))
    x, y = np.array(xs), np.array(ys)
    keep = (x >= -2 * s) & (x < 2 * s) & (y >= -2 * s) & (y < 2 * s)
    x, y = x[keep], y[keep]
    polys = (np.array(part, np.int32), np.array(ring, np.int32), np.array(vx), np.array(vy))
    bbox = (-2 * s, 2 * s, -2 * s, 2 * s)
    for d, max_size in ((6, 32), (3, 256)):
        scale = 4 * s / (1 << d)
        g = run_gpu(cs, x, y, bbox, scale, d, max_size, polys)
        o = run_oracle(oracle_mod, x, y, bbox, scale, d, max_size, polys)
        assert_same(g, o)
        hits = set(zip(g["polygon_index"].tolist(), g["pi"][g["point_index"]].tolist()))
        qi = int(np.flatnonzero((x == quirk[0]) & (y == quirk[1]))[0])
        assert (0, qi) in hits  # polygon 0 is the quirk triangle


def long_ring_scene(seed):
    """Star polygons with long rings (65 to 1,000 edges, some with a 70-edge hole)
    on a grid of disjoint cells, each cell holding a chosen number of points (1 to
    100): on vertices, a few ULPs beside them, on the vertical lines through them,
    and uniform in the cell. With the box filter a polygon's flush then tests
    exactly its own cell's points, so every flush size occurs."""
    rng = np.random.default_rng(seed)
    part, ring, vx, vy, xs, ys = [0], [0], [], [], [], []

    def add_ring(xr, yr):
        vx.extend(list(xr) + [xr[0]])
        vy.extend(list(yr) + [yr[0]])
        ring.append(ring[-1] + len(xr) + 1)

    sizes = [65, 66, 127, 129, 130, 200, 513, 1000]
    counts = [1, 2, 3, 5, 8, 16, 17, 31, 32, 33, 63, 64, 65, 100]
    cell = 8.0
    for c in range(24):
        cx, cy = cell * (c % 6) + cell / 2, cell * (c // 6) + cell / 2
        k = sizes[c % len(sizes)]
        a = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = 3.0 * rng.uniform(0.7, 1.0, k)
        ox, oy = cx + r * np.cos(a), cy + r * np.sin(a)
        add_ring(ox, oy)
        nr = 1
        if c % 3 == 0:  # a hole, clockwise
            h = np.sort(rng.uniform(0, 2 * np.pi, 69))[::-1]
            add_ring(cx + 0.8 * np.cos(h), cy + 0.8 * np.sin(h))
            nr = 2
        part.append(part[-1] + nr)
        m = counts[c % len(counts)]
        vi = rng.integers(0, k, m)
        kind = rng.integers(0, 4, m)
        for v, t in zip(vi, kind):
            if t == 0:
                xs.append(ox[v])
                ys.append(ulps(oy[v], int(rng.integers(-3, 4))))
            elif t == 1:
                xs.append(ulps(ox[v], int(rng.integers(-2, 3))))
                ys.append(oy[v])
            elif t == 2:  # the vertical line through the vertex, inside the cell
                xs.append(ox[v])
                ys.append(cy + rng.uniform(-3.5, 3.5))
            else:
                xs.append(cx + rng.uniform(-3.5, 3.5))
                ys.append(cy + rng.uniform(-3.5, 3.5))
    polys = (np.array(part, np.int32), np.array(ring, np.int32), np.array(vx), np.array(vy))
    return np.array(xs), np.array(ys), polys


@pytest.mark.parametrize("stage,group,box", [(128, 1, 1), (128, 16, 1), (64, 16, 1),
                                             (128, 4096, 0), (128, 0, 1)])
def test_long_rings_few_points_match_oracle(cs, oracle_mod, stage, group, box):
    """Polygon-major flushes of 1 to 64 points, whose edge walks split over copies
    of the points (rings longer than one staged batch, holes, points on vertices
    and on the vertical lines through them): the oracle's pairs bit for bit."""
    from cuspatial_amd import autotune

    x, y, polys = long_ring_scene(31)
    bbox = (0.0, 48.0, 0.0, 48.0)
    saved = autotune.current()
    for d, max_size in ((4, 64), (2, 256)):
        try:
            autotune.apply({"pip_stage_edges": stage, "pip_polygon_group": group,
                            "pip_box_filter": box, "pip_waves_per_block": 2})
            g = run_gpu(cs, x, y, bbox, 48.0 / (1 << d), d, max_size, polys)
        finally:
            autotune.apply(saved)
        o = run_oracle(oracle_mod, x, y, bbox, 48.0 / (1 << d), d, max_size, polys)
        assert_same(g, o)
        assert len(o["polygon_index"]) > 0
