"""The point-in-polygon test's box filter, checked against the oracle on the CPU.

The HIP test skips points strictly outside a polygon's vertex box by a margin
d = 2^-48 (|px| + |py| + B) + 2^-300 in x (B the box's largest |coordinate|), and
points strictly above or below it in y, as certainly outside. The claim is about
the reference's rounded arithmetic (detail/algorithm/is_point_in_polygon.cuh:46-100),
so it is checked with the oracle, which restates that arithmetic: every point the
filter would skip must be outside there. The box rule is restated here from the
kernel (qpip_pip.hip, box_excludes; qpip_edges.hpp, ring boxes).

The margin is needed: a point one ULP left of the box can be inside in the
reference's arithmetic (test_reference_calls_a_point_left_of_the_box_inside).
"""

import numpy as np
import pytest

BOX_COORD_MAX = 2.0 ** 400
BOX_RISE_MIN = 2.0 ** -600


def ring_box(vx, vy):
    """(minx, maxx, miny, maxy) of one closed ring, infinite where the filter may
    not be used (qpip_edges.hpp): a coordinate not finite or above 2^400, or a
    non-zero rise of an effective edge below 2^-600 (every edge of these rings is
    effective: no two consecutive vertices are within 4 ULP of each other)."""
    vx, vy = np.asarray(vx, np.float64), np.asarray(vy, np.float64)
    rise = vy - np.roll(vy, 1)
    ok = (np.all(np.abs(vx) <= BOX_COORD_MAX) and np.all(np.abs(vy) <= BOX_COORD_MAX)
          and np.all((rise == 0) | (np.abs(rise) >= BOX_RISE_MIN)))
    if not ok:
        return (-np.inf, np.inf, -np.inf, np.inf)
    return (vx.min(), vx.max(), vy.min(), vy.max())


def box_excludes(box, px, py):
    """The kernel's rule, vectorised over points."""
    minx, maxx, miny, maxy = box
    px, py = np.asarray(px, np.float64), np.asarray(py, np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        bm = max(abs(minx), abs(maxx), abs(miny), abs(maxy))
        s3 = np.abs(px) + np.abs(py) + bm
        d = s3 * 2.0 ** -48 + 2.0 ** -300  # an fma in the kernel; the margin has room
        out_x = (s3 < 2.0 ** 402) & ((px < minx - d) | (px > maxx + d))
        return (py < miny) | (py > maxy) | out_x


def inside(oracle_mod, vx, vy, px, py):
    """The oracle's answer for one closed ring (one polygon) and many points."""
    n = len(px)
    ring = np.array([0, len(vx)], np.int32)
    bits = oracle_mod.point_in_polygon_bitmask(np.asarray(px, np.float64),
                                               np.asarray(py, np.float64),
                                               np.array([0, 1], np.int32), ring, vx, vy)
    assert len(bits) == n
    return np.asarray(bits) == 1


def quirk_triangle(s=1.0):
    """Leftmost vertex b = (s/2, s/2); the edge from b to a = (s, -s) has
    |run| = s/2 and rise = 3s/2, so for px, py one ULP below s/2 both px - ax and
    py - ay round onto run and rise, the two cross products tie, and that edge does
    not toggle although the point is left of it."""
    vx = np.array([0.5, 1.0, 1.0, 0.5]) * s
    vy = np.array([0.5, -1.0, 1.0, 0.5]) * s
    return vx, vy


def test_reference_calls_a_point_left_of_the_box_inside(oracle_mod):
    for k in (-200, -20, 0, 20, 200):
        s = 2.0 ** k
        vx, vy = quirk_triangle(s)
        px = np.array([np.nextafter(0.5 * s, 0.0)])
        py = px.copy()
        assert px[0] < vx.min()
        assert inside(oracle_mod, vx, vy, px, py)[0], k  # the reference's own rounding
        box = ring_box(vx, vy)
        assert not box_excludes(box, px, py)[0]          # the margin keeps it
        naive = (px < box[0]) | (px > box[1]) | (py < box[2]) | (py > box[3])
        assert naive[0]                                   # a plain box test would not


def _probe_points(box, vx, vy, rng, n_rand=200):
    """Points on all four sides of the box: a few ULPs out, around the margin d,
    and beyond; with y at vertex heights (and a few ULPs off) and at random."""
    minx, maxx, miny, maxy = box
    ys = np.concatenate([vy, [np.nextafter(v, -np.inf) for v in vy],
                         [np.nextafter(v, np.inf) for v in vy],
                         rng.uniform(miny, maxy, n_rand)])
    xs_side = []
    bm = max(abs(minx), abs(maxx), abs(miny), abs(maxy))
    for edge, direction in ((minx, -np.inf), (maxx, np.inf)):
        v = edge
        for _ in range(40):  # 1 .. 40 ULPs out
            v = np.nextafter(v, direction)
            xs_side.append(v)
        d = (bm + abs(edge) + np.abs(ys).max()) * 2.0 ** -48
        sign = -1.0 if direction < 0 else 1.0
        for f in (0.5, 0.9, 0.99, 1.0, 1.01, 1.1, 2.0, 10.0, 1e3):
            xs_side.append(edge + sign * f * d)
    xs_side = np.array(xs_side)
    px = np.repeat(xs_side, len(ys))
    py = np.tile(ys, len(xs_side))
    # above and below: x across the box
    xs = np.concatenate([vx, rng.uniform(minx, maxx, 50)])
    for edge, direction in ((miny, -np.inf), (maxy, np.inf)):
        v = edge
        for _ in range(4):
            v = np.nextafter(v, direction)
            px = np.concatenate([px, xs])
            py = np.concatenate([py, np.full(len(xs), v)])
    return px, py


def _shapes(rng, s):
    """Closed rings (first vertex repeated) with long nearly vertical and nearly
    horizontal edges ending at the box's extremes."""
    out = [quirk_triangle(s)]
    for _ in range(6):
        k = int(rng.integers(3, 9))
        a = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = rng.uniform(0.3, 1.0, k)
        vx = np.append(s * r * np.cos(a), s * r[0] * np.cos(a[0]))
        vy = np.append(s * r * np.sin(a), s * r[0] * np.sin(a[0]))
        out.append((vx, vy))
    for _ in range(4):  # slivers: one long edge a few ULPs from vertical / horizontal
        x0 = s * rng.uniform(-1, 1)
        y0 = s * rng.uniform(-1, 1)
        x1 = x0
        for _ in range(int(rng.integers(1, 6))):
            x1 = np.nextafter(x1, np.inf)
        out.append((np.array([x0, x1, x0 - s * 0.3, x0]), np.array([y0, y0 + s, y0 + 0.5 * s, y0])))
        y1 = y0
        for _ in range(int(rng.integers(1, 6))):
            y1 = np.nextafter(y1, np.inf)
        out.append((np.array([x0, x0 + s, x0 + 0.5 * s, x0]), np.array([y0, y1, y0 - 0.3 * s, y0])))
    return out


@pytest.mark.parametrize("k", [-700, -250, -30, 0, 7, 40, 250, 390])
def test_filtered_points_are_outside_in_the_reference(oracle_mod, k):
    """Every probe point the filter skips is outside in the oracle, at scales from
    2^-700 (rises below 2^-600: the box is infinite and skips nothing) to 2^390."""
    rng = np.random.default_rng(1000 + k)
    s = 2.0 ** k
    skipped = 0
    for vx, vy in _shapes(rng, s):
        box = ring_box(vx, vy)
        px, py = _probe_points(box if np.isfinite(box[0]) else
                               (vx.min(), vx.max(), vy.min(), vy.max()), vx, vy, rng)
        ex = box_excludes(box, px, py)
        got = inside(oracle_mod, vx, vy, px[ex], py[ex])
        assert not got.any(), (k, px[ex][got][:3], py[ex][got][:3])
        skipped += int(ex.sum())
    if k >= -250:
        assert skipped > 1000  # the filter is exercised
    else:
        assert skipped == 0    # tiny rises: nothing is skipped


def test_margin_is_not_dead_weight(oracle_mod):
    """Across random quirk-prone triangles, points a few ULPs left of the box that
    the oracle calls inside exist, and the filter keeps every one of them."""
    rng = np.random.default_rng(5)
    found = 0
    for _ in range(300):
        s = 2.0 ** int(rng.integers(-20, 20))
        bx, by = 0.5 * s, 0.5 * s
        ax, ay = s * rng.uniform(0.6, 1.0), -s * rng.uniform(0.6, 1.0)
        vx = np.array([bx, ax, ax, bx])
        vy = np.array([by, ay, s, by])
        box = ring_box(vx, vy)
        px, py = [], []
        x = bx
        for _ in range(16):
            x = np.nextafter(x, -np.inf)
            y = by
            for _ in range(16):
                y = np.nextafter(y, -np.inf)
                px.append(x)
                py.append(y)
        px, py = np.array(px), np.array(py)
        got = inside(oracle_mod, vx, vy, px, py)
        found += int(got.sum())
        assert not (got & box_excludes(box, px, py)).any()
    assert found > 0


def test_empty_and_nan_rings():
    """No vertices: every point is outside (nothing toggles), and the empty box
    skips finite points. A NaN vertex: the box is infinite and skips nothing."""
    inf = np.inf
    empty = (inf, -inf, inf, -inf)
    assert box_excludes(empty, np.array([0.0, 1e300]), np.array([0.0, -1e300])).all()
    assert not box_excludes(empty, np.array([np.nan]), np.array([np.nan])).any()
    box = ring_box([0.0, 1.0, np.nan, 0.0], [0.0, 0.0, 1.0, 0.0])
    assert box == (-inf, inf, -inf, inf)
    assert not box_excludes(box, np.array([5.0, -5.0]), np.array([5.0, 0.5])).any()
