"""contains_properly on the GPU path (python/cuspatial/cuspatial/core/binpreds/
contains.py, feature_contains_properly.py, contains_geometry_processor.py).

Bar: the reference's own expectations (tests/golden/contains_properly_kats.json,
from tests/binpreds/test_contains_properly.py) exactly, and, for larger inputs,
pair sets equal to the oracle's quadtree pipeline run with the caller's
parameters and remapped the same way.
"""

from math import ceil, sqrt

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


def polys(cs, rings_per_polygon, geometry_offset=None):
    part, ring, verts = [0], [0], []
    for rings in rings_per_polygon:
        for r in rings:
            verts += r
            ring.append(len(verts))
        part.append(len(ring) - 1)
    v = np.array(verts, dtype=np.float64).reshape(-1, 2)
    return cs.Polygons(part, ring, v[:, 0], v[:, 1], geometry_offset=geometry_offset,
                       device="cuda")


def pts(cs, xy):
    a = np.asarray(xy, dtype=np.float64).reshape(-1, 2)
    return cs.Points(a[:, 0], a[:, 1], device="cuda")


@pytest.mark.parametrize("case", load_golden("contains_properly_kats.json"),
                         ids=lambda c: "%s_%s" % (c["source"].split(":")[-1], c["points"][0]))
def test_contains_properly_kats(cs, case):
    from cuspatial_amd.binpreds import polygon_contains_properly

    P = polys(cs, [[r] for r in case["polygons"]])
    got = polygon_contains_properly(P, pts(cs, case["points"])).cpu().numpy()
    assert got.tolist() == case["expected"]


@pytest.mark.parametrize("mode", ["quadtree", "pairwise", "brute_force"])
def test_modes_agree_on_kats(cs, mode):
    """Each of contains.py's three modes gives the same containments for the
    single point/polygon vectors (mode-specific column layouts)."""
    from cuspatial_amd.binpreds import contains_properly

    for case in load_golden("contains_properly_kats.json"):
        if len(case["points"]) != 1 or len(case["polygons"]) != 1:
            continue
        r = contains_properly(polys(cs, [[case["polygons"][0]]]), pts(cs, case["points"]),
                              mode=mode)
        if mode == "quadtree":
            assert r.columns in ([], ["point_index", "part_index"])
        elif mode == "pairwise":
            assert r.columns == ["pairwise_index", "point_index", "result"]
        else:
            assert r.columns == ["point_index", "part_index", "result"]
        assert (len(r) == 1) == case["expected"][0], case


def oracle_quadtree_pairs(oracle, x, y, part, ring, vx, vy):
    """contains.py:19-74 on the oracle: same parameters, same remap."""
    d = 15
    min_size = ceil(sqrt(len(x)))
    x0, x1, y0, y1 = vx.min(), vx.max(), vy.min(), vy.max()
    scale = max(x1 - x0, y1 - y0) / ((1 << d) + 2)
    pi, key, level, internal, length, offset = oracle.quadtree_on_points(x, y, x0, x1, y0, y1,
                                                                          scale, d, min_size)
    tree = (key, level, internal, length, offset)
    bb = oracle.polygon_bounding_boxes(part, ring, vx, vy)
    jb, jq = oracle.join_quadtree_and_bounding_boxes(tree, bb, x0, x1, y0, y1, scale, d)
    pp, pt = oracle.quadtree_point_in_polygon((jb, jq), tree, pi, x, y, part, ring, vx, vy)
    return np.asarray(pi)[np.asarray(pt, np.int64)].astype(np.int64), np.asarray(pp, np.int64)


def test_quadtree_mode_matches_oracle(cs, oracle_mod):
    from cuspatial_amd.binpreds import contains_properly

    rng = np.random.default_rng(21)
    part, ring, vx, vy = star_polygons(rng, 60, 1000.0, 20.0, 120.0, holes=0.3)
    x, y = rng.uniform(0, 1000, 30000), rng.uniform(0, 1000, 30000)
    P = cs.Polygons(part, ring, vx, vy, device="cuda")
    r = contains_properly(P, cs.Points(x, y, device="cuda"), mode="quadtree")
    opt, opp = oracle_quadtree_pairs(oracle_mod, x, y, part, ring, vx, vy)
    assert r.columns == ["point_index", "part_index"]
    assert r["point_index"].cpu().numpy().astype(np.int64).tolist() == opt.tolist()
    assert r["part_index"].cpu().numpy().astype(np.int64).tolist() == opp.tolist()
    assert len(opt) > 1000


def test_allpairs_and_pairwise_modes(cs, oracle_mod):
    """More than 31 polygons: allpairs=True takes the quadtree mode and returns
    the deduplicated (polygon_index, point_index) set; allpairs=False with as
    many points as polygons takes the pairwise mode (point i against polygon i)."""
    from cuspatial_amd.binpreds import pip_mode, polygon_contains_properly

    rng = np.random.default_rng(22)
    part, ring, vx, vy = star_polygons(rng, 48, 100.0, 5.0, 30.0, holes=0.3)
    n = len(part) - 1
    x, y = rng.uniform(0, 100, n), rng.uniform(0, 100, n)
    P = cs.Polygons(part, ring, vx, vy, device="cuda")
    Q = cs.Points(x, y, device="cuda")
    assert pip_mode(P, allpairs=True) == "quadtree" and pip_mode(P) == "pairwise"
    ap = polygon_contains_properly(P, Q, allpairs=True)
    opt, opp = oracle_quadtree_pairs(oracle_mod, x, y, part, ring, vx, vy)
    want = sorted(set(zip(opp.tolist(), opt.tolist())))
    got = list(zip(ap["polygon_index"].cpu().tolist(), ap["point_index"].cpu().tolist()))
    assert got == want
    single = polygon_contains_properly(P, Q).cpu().numpy()
    ow = oracle_mod.pairwise_point_in_polygon(x, y, part, ring, vx, vy)
    assert single.tolist() == [bool(v) for v in ow]


def test_multipolygon_parts_map_to_geometries(cs, oracle_mod):
    """Parts grouped into multipolygons by geometry_offset: the quadtree mode's
    part hits become geometry hits, deduplicated (a point can be inside two parts
    of one geometry only if the parts overlap, as they may here)."""
    from cuspatial_amd.binpreds import polygon_contains_properly

    rng = np.random.default_rng(23)
    part, ring, vx, vy = star_polygons(rng, 70, 100.0, 5.0, 25.0, holes=0.2)
    geom = np.array([0, 3, 4, 10] + list(range(11, 71)), dtype=np.int32)
    assert geom[-1] == 70
    x, y = rng.uniform(0, 100, 5000), rng.uniform(0, 100, 5000)
    P = cs.Polygons(part, ring, vx, vy, geometry_offset=geom, device="cuda")
    ap = polygon_contains_properly(P, cs.Points(x, y, device="cuda"), allpairs=True)
    opt, opp = oracle_quadtree_pairs(oracle_mod, x, y, part, ring, vx, vy)
    og = np.searchsorted(geom, opp, side="right") - 1
    want = sorted(set(zip(og.tolist(), opt.tolist())))
    got = list(zip(ap["polygon_index"].cpu().tolist(), ap["point_index"].cpu().tolist()))
    assert got == want and len(want) > 100


def test_index_labels_and_unequal_lengths(cs):
    """The relabel through the left series' index, and the unequal-length point
    result (contains_geometry_processor.py:273-275)."""
    from cuspatial_amd.binpreds import polygon_contains_properly

    sq = [[0.0, 0.0], [4.0, 0.0], [4.0, 4.0], [0.0, 4.0], [0.0, 0.0]]
    far = [[10.0, 10.0], [14.0, 10.0], [14.0, 14.0], [10.0, 14.0], [10.0, 10.0]]
    P = polys(cs, [[sq], [far]])
    Q = pts(cs, [[1.0, 1.0], [11.0, 11.0], [20.0, 20.0]])
    ap = polygon_contains_properly(P, Q, allpairs=True, index=[7, 9])
    assert ap["polygon_index"].cpu().tolist() == [7, 9]
    assert ap["point_index"].cpu().tolist() == [0, 1]
    # two polygons, three points: point j is set for every polygon position j with a hit
    got = polygon_contains_properly(P, Q).cpu().tolist()
    assert got == [True, True, False]
