"""Pins of the ranker oracle (CPU): its gradient against finite differences, the
bf16 rounding against the format's definition, Adam against its update rule,
and the candidate grid against brute force."""

import numpy as np
import pytest

from conftest import star_polygons


@pytest.fixture(scope="module")
def rk():
    from oracle import ranker

    return ranker


def small_problem(rk, seed=0, cells=9, n=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, rk.layout(cells)["total"])
    X = np.concatenate([rng.normal(0, 1, (n, rk.EMB)), rng.uniform(-1, 1, (n, rk.FOUR))], 1)
    pc = rng.integers(0, cells, n)
    counts = rng.integers(0, 18, cells)
    labels = np.where(rng.uniform(size=n) < 0.8, rng.integers(0, 16, n), -1)
    return p, X, pc, counts, labels


def test_gradient_matches_finite_differences(rk):
    """float64, no rounding: every parameter block's analytic gradient equals
    central differences of the loss (the embedding enters through X = [E[cell], F])."""
    cells, n = 9, 40
    p, X, pc, counts, labels = small_problem(rk, cells=cells, n=n)
    Xf = X[:, rk.EMB:]

    def loss_of(q):
        E = rk.unpack(q, cells)["E"]
        XX = np.concatenate([E[pc], Xf], 1)
        return rk.loss_and_grad(q, cells, XX, pc, counts, labels, bf16=False)[0]

    E = rk.unpack(p, cells)["E"]
    _, g, st = rk.loss_and_grad(p, cells, np.concatenate([E[pc], Xf], 1), pc, counts, labels,
                                bf16=False)
    assert st["labelled"] > 10
    rng = np.random.default_rng(1)
    o = rk.layout(cells)
    for a, b in (("e", "w1"), ("w1", "b1"), ("b1", "w2"), ("w2", "b2"), ("b2", "w3"),
                 ("w3", "b3"), ("b3", "total")):
        idx = rng.integers(o[a], o[b], 12)
        for i in idx:
            h = 1e-6
            q1, q2 = p.copy(), p.copy()
            q1[i] += h
            q2[i] -= h
            fd = (loss_of(q1) - loss_of(q2)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (a, i, fd, g[i])


def test_bf16_round_definition(rk):
    """Round to nearest, ties to even, on the 16 dropped bits; NaN stays NaN."""
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0, -0.0,
                     3.4e38, 1e-40, np.inf, -np.inf], np.float32)
    r = rk.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == np.float32(1.0)  # 1 + 2^-8 is a tie: to the even mantissa (1.0)
    assert r[2] == np.float32(1.0)  # below half
    assert r[3] == np.float32(1.0 + 2 ** -7)  # above half
    assert r[4] == np.float32(-2.5) and r[5] == 0 and np.signbit(r[6])
    assert np.isinf(r[9]) and np.isinf(r[10])
    assert np.isnan(rk.bf16_round(np.array([np.nan], np.float32)))[0]
    # 1 + 3 * 2^-8 is a tie between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    assert rk.bf16_round(np.array([1.0 + 3 * 2 ** -8], np.float32))[0] == np.float32(1.0 + 2 ** -6)
    # every value is representable in 8 mantissa bits after rounding
    x = np.random.default_rng(0).normal(0, 10, 10000).astype(np.float32)
    assert np.all((rk.bf16_round(x).view(np.uint32) & 0xFFFF) == 0)
    assert np.max(np.abs(rk.bf16_round(x) - x) / np.abs(x)) <= 2 ** -8


def test_adam_rule(rk):
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.1, -0.3, 0.0])
    q, m, v = rk.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    # the first step moves every parameter with a nonzero gradient by lr against its sign
    assert np.allclose(q, p - 0.01 * np.sign(g), atol=1e-9)


def test_candidate_grid_is_complete(rk, oracle_mod):
    """Every polygon that contains a point is listed in the point's cell."""
    rng = np.random.default_rng(5)
    part, ring, vx, vy = star_polygons(rng, 80, 100.0, 2.0, 12.0, holes=0.3)
    x0, y0, dim = 0.0, 0.0, 16
    cell = 100.0 / dim
    lists = rk.candidate_lists(part, ring, vx, vy, x0, y0, cell, dim)
    x = rng.uniform(-5, 105, 20000)
    y = rng.uniform(-5, 105, 20000)
    pi, key, level, internal, length, offset = oracle_mod.quadtree_on_points(
        x, y, -5, 105, -5, 105, 110 / 64, 6, 50)
    tree = (key, level, internal, length, offset)
    bb = oracle_mod.polygon_bounding_boxes(part, ring, vx, vy)
    jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(tree, bb, -5, 105, -5, 105, 110 / 64, 6)
    pp, pt = oracle_mod.quadtree_point_in_polygon((jb, jq), tree, pi, x, y, part, ring, vx, vy)
    pts = np.asarray(pi)[np.asarray(pt, np.int64)]
    cells = rk.point_cells(x[pts], y[pts], x0, y0, cell, dim)
    assert len(pts) > 1000
    for c, poly in zip(cells, np.asarray(pp)):
        assert poly in lists.get(int(c), [])


def test_grid_builder_matches_brute_force(rk):
    """CandidateGrid.build (host code) lists exactly the oracle's polygons per
    cell, in ascending id, and marks cells with more than 16 as overflowed."""
    from cuspatial_amd import Polygons
    from cuspatial_amd.ranker import OVERFLOW, CandidateGrid

    rng = np.random.default_rng(8)
    part, ring, vx, vy = star_polygons(rng, 300, 100.0, 2.0, 15.0, holes=0.3)
    part = np.concatenate([part[:10], [part[9]], part[10:] + 0])  # an empty polygon at 9
    part = np.asarray(part, np.int32)
    P = Polygons(part, ring, vx, vy, device="cpu")
    dim = 12
    g = CandidateGrid.build(P, 0.0, 0.0, 100.0, dim)
    want = rk.candidate_lists(part, ring, vx, vy, 0.0, 0.0, 100.0 / dim, dim)
    cand = g.candidates.numpy().reshape(-1, 16)
    counts = g.counts.numpy()
    assert (counts == OVERFLOW).any() and (counts < 16).any()
    for c in range(dim * dim):
        lst = want.get(c, [])
        if len(lst) > 16:
            assert counts[c] == OVERFLOW
        else:
            assert counts[c] == len(lst)
            assert cand[c, :len(lst)].tolist() == lst and (cand[c, len(lst):] == -1).all()
