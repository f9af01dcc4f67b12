"""quadtree_point_to_nearest_linestring and linestring_bounding_boxes on the GPU.

Bar: bit-identical to the reference's expected vectors
(tests/golden/nearest_linestring_small.json) and to the oracle on random inputs
(indices exactly; distances bit for bit, the same fma-contracted arithmetic).
"""

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cs():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd

    return cuspatial_amd


def np_(t):
    return t.cpu().numpy()


def run_gpu(cs, x, y, offs, lx, ly, bbox, scale, d, ms, radius):
    pts = cs.Points(x, y, device="cuda")
    L = cs.Linestrings(offs, lx, ly, device="cuda")
    pi, tree = cs.quadtree_on_points(pts, *bbox, scale, d, ms)
    bb = cs.linestring_bounding_boxes(L, radius)
    pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *bbox, scale, d)
    res = cs.quadtree_point_to_nearest_linestring(pairs, tree, pi, pts, L)
    return pi, tree, bb, pairs, res


def test_golden_small(cs):
    g = load_golden("nearest_linestring_small.json")
    p = np.array(g["points"])
    v = np.array(g["vertices"])
    pi, tree, bb, pairs, res = run_gpu(cs, p[:, 0], p[:, 1], g["linestring_offsets"], v[:, 0],
                                       v[:, 1], g["bbox"], g["scale"], g["max_depth"],
                                       g["max_size"], g["expansion_radius"])
    assert np_(pairs["bbox_offset"]).tolist() == g["expected_join_linestring_index"]
    assert np_(pairs["quad_offset"]).tolist() == g["expected_join_quad_index"]
    assert np_(res["point_index"]).tolist() == g["expected_point_index"]
    assert np_(res["linestring_index"]).tolist() == g["expected_linestring_index"]
    assert np.array_equal(np_(res["distance"]).view(np.int64),
                          np.array(g["expected_distance"]).view(np.int64))


def random_lines(rng, L, ext, kmin=2, kmax=7, spread=0.06):
    offs, lx, ly = [0], [], []
    for _ in range(L):
        k = int(rng.integers(kmin, kmax + 1))
        cx, cy = rng.uniform(0, ext, 2)
        lx += list(cx + rng.normal(0, spread * ext, k))
        ly += list(cy + rng.normal(0, spread * ext, k))
        offs.append(len(lx))
    return np.array(offs, np.int32), np.array(lx), np.array(ly)


@pytest.mark.parametrize("seed,n,L,d,ms,radius", [
    (1, 20000, 50, 6, 64, 2.0),      # many leaves, few linestrings each
    (2, 5000, 400, 3, 2000, 5.0),    # few large leaves, long runs of pairs
    (3, 30000, 10, 8, 16, 0.5),      # uncovered leaves (small radius)
    (4, 100000, 200, 10, 100, 1.0),
])
def test_random_matches_oracle(cs, oracle_mod, seed, n, L, d, ms, radius):
    rng = np.random.default_rng(seed)
    ext = 100.0
    x, y = rng.uniform(0, ext, n), rng.uniform(0, ext, n)
    offs, lx, ly = random_lines(rng, L, ext)
    bbox = (0.0, ext, 0.0, ext)
    scale = ext / (1 << d)
    pi, tree, bb, pairs, res = run_gpu(cs, x, y, offs, lx, ly, bbox, scale, d, ms, radius)
    obb = oracle_mod.linestring_bounding_boxes(offs, lx, ly, radius)
    for c, o in zip(("minx", "miny", "maxx", "maxy"), obb):
        assert np.array_equal(np_(bb[c]), o)
    t = tuple(np_(tree[k]) for k in ("key", "level", "is_internal_node", "length", "offset"))
    opt, ols, od = oracle_mod.quadtree_point_to_nearest_linestring(
        (np_(pairs["bbox_offset"]), np_(pairs["quad_offset"])), t, np_(pi), x, y, offs, lx, ly)
    assert np.array_equal(np_(res["point_index"]), opt)
    assert np.array_equal(np_(res["linestring_index"]), ols)
    assert np.array_equal(np_(res["distance"]).view(np.int64), od.view(np.int64))
    covered = int((ols != 0xFFFFFFFF).sum())
    assert covered > 0
    if seed == 3:
        assert covered < n  # some points lie in leaves with no pair


def test_empty_and_errors(cs):
    import torch

    rng = np.random.default_rng(5)
    offs, lx, ly = random_lines(rng, 5, 10.0)
    x, y = rng.uniform(0, 10, 100), rng.uniform(0, 10, 100)
    pi, tree, bb, pairs, res = run_gpu(cs, x, y, offs, lx, ly, (0.0, 10.0, 0.0, 10.0), 10 / 8, 3,
                                       10, 1.0)
    L = cs.Linestrings(offs, lx, ly, device="cuda")
    # empty point data (quadtree_point_to_nearest_linestring_test.cpp:149-174)
    e = cs.Points(np.zeros(0), np.zeros(0), device="cuda")
    r = cs.quadtree_point_to_nearest_linestring(pairs, tree, torch.zeros(0, dtype=torch.int32,
                                                                        device="cuda"), e, L)
    assert all(len(v) == 0 for v in r.values()) and r.columns == ["point_index",
                                                                  "linestring_index", "distance"]
    # empty linestring data (:176-199)
    r = cs.quadtree_point_to_nearest_linestring(pairs, tree, pi, cs.Points(x, y, device="cuda"),
                                                cs.Linestrings(np.zeros(0, np.int32), np.zeros(0),
                                                               np.zeros(0), device="cuda"))
    assert all(len(v) == 0 for v in r.values())
    with pytest.raises(RuntimeError, match="expansion radius must be greater or equal than 0"):
        cs.linestring_bounding_boxes(L, -1.0)
    with pytest.raises(RuntimeError, match="all linestrings must have at least 2 vertices"):
        cs.linestring_bounding_boxes(cs.Linestrings([0, 1, 2], [0.0, 1.0], [0.0, 1.0],
                                                    device="cuda"), 0.0)
    with pytest.raises(ValueError, match="cannot contain multilinestrings"):
        cs.quadtree_point_to_nearest_linestring(
            pairs, tree, pi, cs.Points(x, y, device="cuda"),
            cs.Linestrings(offs, lx, ly, geometry_offset=[0, 2, 5], device="cuda"))
