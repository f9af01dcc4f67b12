"""The learned ranker on the GPU against its oracle (oracle/ranker.py) and the
learned locator against the plain quadtree join.

Bars (stated per test): features within one bf16 step of the oracle's; logits
and gradients within a relative norm error of 2e-2 of the oracle's bf16
emulation (fp32 MFMA accumulation against float64, activations rounded to bf16
at the same points); Adam within 1e-6 of its rule; the locator's (polygon,
point) pairs bit-identical to the plain join's.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd as cs
    from cuspatial_amd import datasets, ranker
    from oracle import ranker as rk

    return torch, cs, ranker, rk, datasets


def voronoi(env, n_points, n_seeds, seed, dim):
    torch, cs, ranker, rk, datasets = env
    part, ring, px, py = datasets.voronoi_cells(seed, n_seeds, 1e5)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1e5, n_points)
    y = rng.uniform(0, 1e5, n_points)
    xy = torch.as_tensor(np.stack([x, y], 1).reshape(-1), device="cuda")
    P = cs.Polygons(part, ring, px, py, device="cuda")
    grid = ranker.CandidateGrid.build(P, 0.0, 0.0, 1e5, dim)
    return (part, ring, px, py), P, grid, xy, x, y


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_features_match_oracle(env):
    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 5000, 200, 3, 16)
    model = ranker.Ranker(grid.num_cells, seed=1)
    X, cells = model.features(grid, xy)
    p = model.params.cpu().numpy().astype(np.float64)
    Xo, co = rk.features(p, grid.num_cells, x, y, 0.0, 0.0, grid.cell, grid.dim)
    assert np.array_equal(cells.cpu().numpy().astype(np.int64), co)
    Xg = rk.from_bf16_bits(X.cpu().numpy().view(np.uint16))
    assert np.array_equal(Xg[:, :64], Xo[:, :64].astype(np.float32))  # embedding rows, exact
    # Fourier features: GPU sinf/cosf against numpy, then bf16: one bf16 step at most
    assert np.max(np.abs(Xg[:, 64:] - Xo[:, 64:])) <= 2 ** -8 * 1.01


def test_logits_match_oracle(env):
    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 3001, 200, 4, 16)  # not a multiple of 16: padding
    model = ranker.Ranker(grid.num_cells, seed=2)
    z = model.logits(grid, xy).cpu().numpy()
    X, _ = model.features(grid, xy)
    Xg = rk.from_bf16_bits(X.cpu().numpy().view(np.uint16)).astype(np.float64)
    zo = rk.forward(model.params.cpu().numpy().astype(np.float64), grid.num_cells, Xg)["Z"]
    assert relerr(z, zo) < 2e-2, relerr(z, zo)
    # the MFMA lane layout: a wrong operand order or transpose would be O(1) off
    assert np.corrcoef(z.reshape(-1), zo.reshape(-1))[0, 1] > 0.999


def test_gradients_match_oracle(env):
    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 2000, 150, 5, 16)
    model = ranker.Ranker(grid.num_cells, seed=3)
    rng = np.random.default_rng(6)
    labels = np.where(rng.uniform(size=2000) < 0.9, rng.integers(0, 6, 2000), -1).astype(np.int32)
    grad, stats = model.gradients(grid, xy, torch.as_tensor(labels, device="cuda"))
    g = grad.cpu().numpy()
    st = stats.cpu().numpy()
    X, cells = model.features(grid, xy)
    Xg = rk.from_bf16_bits(X.cpu().numpy().view(np.uint16)).astype(np.float64)
    loss, go, so = rk.loss_and_grad(model.params.cpu().numpy().astype(np.float64),
                                    grid.num_cells, Xg, cells.cpu().numpy().astype(np.int64),
                                    grid.counts.cpu().numpy(), labels)
    assert so["labelled"] > 1000 and st[2] == so["labelled"]
    assert abs(st[0] - so["loss_sum"]) <= 1e-3 * so["loss_sum"]
    o = rk.layout(grid.num_cells)
    for a, b in (("e", "w1"), ("w1", "b1"), ("b1", "w2"), ("w2", "b2"), ("b2", "w3"),
                 ("w3", "b3"), ("b3", "total")):
        assert relerr(g[o[a]:o[b]], go[o[a]:o[b]]) < 2e-2, (a, relerr(g[o[a]:o[b]], go[o[a]:o[b]]))


def test_adam_step_matches_rule(env):
    torch, cs, ranker, rk, _ = env
    model = ranker.Ranker(16, seed=4)
    p0 = model.params.cpu().numpy().astype(np.float64)
    g = torch.randn(model.params.numel(), generator=torch.Generator().manual_seed(1)) * 0.01
    model.grad.copy_(g.to("cuda"))
    model.adam(1e-2)
    model.grad.copy_((g * 2).to("cuda"))
    model.adam(1e-2)
    q, m, v = rk.adam(p0, g.numpy().astype(np.float64), 0.0, 0.0, 1, 1e-2)
    q, m, v = rk.adam(q, 2 * g.numpy().astype(np.float64), m, v, 2, 1e-2)
    got = model.params.cpu().numpy()
    assert np.max(np.abs(got - q)) < 1e-6
    assert torch.equal(model.params_bf16.cpu(), model.params.to(torch.bfloat16).cpu())


def exact_pairs(env, P, xy, bbox=(0.0, 1e5, 0.0, 1e5), d=10, ms=64):
    torch, cs, ranker, rk, _ = env
    pts = cs.Points.from_points_xy(xy)
    return ranker.plain_pairs(pts, P, bbox, (bbox[1] - bbox[0]) / (1 << d), d, ms)


def test_locator_equals_plain_join_untrained_and_trained(env):
    """Exact whatever the model: an untrained model, a zero model (list order),
    and a trained one all give the plain join's pairs; training lowers the tests
    per point below list order's."""
    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 200000, 400, 7, 32)
    want_poly, want_pt = exact_pairs(env, P, xy)
    per_point = ranker.per_point_polygon(want_poly, want_pt, len(x))
    tests = {}
    for name, model in (("random", ranker.Ranker(grid.num_cells, seed=5)),
                        ("zero", ranker.Ranker(grid.num_cells, zero=True))):
        loc = ranker.LearnedLocator(P, grid, model, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
        res = loc.locate(cs.Points.from_points_xy(xy))
        poly, pt = res.pairs()
        assert torch.equal(poly, want_poly) and torch.equal(pt, want_pt), name
        tests[name] = res.tests.to(torch.float64).mean().item()
    labels = grid.slots_of(xy, per_point)
    model = ranker.Ranker(grid.num_cells, seed=6)
    hist = model.train(grid, xy, labels, epochs=6, batch=8192, lr=5e-3)
    assert hist[-1][0] < hist[0][0] and hist[-1][1] > hist[0][1], hist
    loc = ranker.LearnedLocator(P, grid, model, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    res = loc.locate(cs.Points.from_points_xy(xy))
    poly, pt = res.pairs()
    assert torch.equal(poly, want_poly) and torch.equal(pt, want_pt)
    trained = res.tests.to(torch.float64).mean().item()
    assert trained < tests["zero"], (trained, tests)


def test_locator_fallback_paths(env):
    """max_tests = 1 sends every first-prediction miss to the quadtree path; an
    overflowed cell (more than 16 candidates) sends all its points there; the
    answer stays exact."""
    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 50000, 3000, 8, 16)  # coarse grid: some cells overflow
    assert int((grid.counts == ranker.OVERFLOW).sum()) > 0
    want_poly, want_pt = exact_pairs(env, P, xy)
    model = ranker.Ranker(grid.num_cells, seed=9)
    loc = ranker.LearnedLocator(P, grid, model, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    for mt in (1, 16):
        res = loc.locate(cs.Points.from_points_xy(xy), max_tests=mt)
        poly, pt = res.pairs()
        assert torch.equal(poly, want_poly) and torch.equal(pt, want_pt), mt
        assert res.fallback > 0
    # points outside the domain and on a shared vertex take the exact answer too
    part, ring, px, py = (P.part_offset.cpu().numpy(), P.ring_offset.cpu().numpy(),
                          P.x.cpu().numpy(), P.y.cpu().numpy())
    extra = np.array([[-10.0, 5.0], [1e5 + 1, 3.0], [px[3], py[3]], [np.nan, 1.0]])
    xy2 = torch.as_tensor(extra.reshape(-1), device="cuda")
    res = loc.locate(cs.Points.from_points_xy(xy2))
    wp, wt = exact_pairs(env, P, xy2)
    poly, pt = res.pairs()
    assert torch.equal(poly, wp) and torch.equal(pt, wt)


def test_locator_errors(env):
    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 100, 50, 10, 8)
    model = ranker.Ranker(16 * 16)  # embeddings for a 16 x 16 grid, grid is 8 x 8
    loc = ranker.LearnedLocator(P, grid, model, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    with pytest.raises(RuntimeError, match="embedding count must equal"):
        loc.locate(cs.Points.from_points_xy(xy))
    model = ranker.Ranker(grid.num_cells)
    loc = ranker.LearnedLocator(P, grid, model, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    with pytest.raises(RuntimeError, match="max_tests must be at least 1"):
        loc.locate(cs.Points.from_points_xy(xy), max_tests=0)


def test_locator_partition_with_holes(env):
    """A partition whose first 40 cells have a hole (the cell shrunk towards its
    centroid) filled by an island polygon: the holed cells take the ring-table
    path of the exact test, the islands the single-ring path; pairs equal the
    plain join's."""
    torch, cs, ranker, rk, datasets = env
    part, ring, px, py = datasets.voronoi_cells(12, 300, 1e5)
    parts, rings, xs, ys = [0], [0], [], []
    islands = []
    for i in range(len(part) - 1):
        a, b = ring[part[i]], ring[part[i + 1]]
        vx, vy = px[a:b], py[a:b]
        xs.append(vx)
        ys.append(vy)
        rings.append(rings[-1] + len(vx))
        nr = 1
        if i < 40:
            cx, cy = vx[:-1].mean(), vy[:-1].mean()
            hx, hy = cx + 0.4 * (vx - cx), cy + 0.4 * (vy - cy)
            xs.append(hx)
            ys.append(hy)
            rings.append(rings[-1] + len(hx))
            nr = 2
            islands.append((hx, hy))
        parts.append(parts[-1] + nr)
    for hx, hy in islands:
        xs.append(hx)
        ys.append(hy)
        rings.append(rings[-1] + len(hx))
        parts.append(parts[-1] + 1)
    P = cs.Polygons(np.array(parts, np.int32), np.array(rings, np.int32), np.concatenate(xs),
                    np.concatenate(ys), device="cuda")
    grid = ranker.CandidateGrid.build(P, 0.0, 0.0, 1e5, 32)
    rng = np.random.default_rng(13)
    x, y = rng.uniform(0, 1e5, 100000), rng.uniform(0, 1e5, 100000)
    xy = torch.as_tensor(np.stack([x, y], 1).reshape(-1), device="cuda")
    want_poly, want_pt = exact_pairs(env, P, xy)
    assert int((want_poly >= 300).sum()) > 100  # points inside islands
    loc = ranker.LearnedLocator(P, grid, ranker.Ranker(grid.num_cells, seed=14),
                                (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    res = loc.locate(cs.Points.from_points_xy(xy))
    poly, pt = res.pairs()
    assert torch.equal(poly, want_poly) and torch.equal(pt, want_pt)


def test_fused_ranking_matches_gemm_logits(env):
    """The locator's fused MLP (rank_kernel, 16x16x32 MFMA, weights chained in
    registers) ranks like the training path's logits (qpip_ranker_logits, checked
    against the oracle above): its first prediction is the argmax of those logits
    over the cell's candidates, up to near-ties (accumulation order differs)."""
    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 100000, 400, 15, 32)
    model = ranker.Ranker(grid.num_cells, seed=16)
    z = model.logits(grid, xy)
    c = grid.cells_of(xy)
    cnt = grid.counts[c].to(torch.int64)
    mask = torch.arange(16, device="cuda")[None, :] < cnt[:, None]
    zm = torch.where(mask, z, torch.full_like(z, -float("inf")))
    top = zm.argmax(1)
    srt = zm.sort(1, descending=True).values
    clear = (srt[:, 0] - srt[:, 1]) > 1e-3 * srt[:, 0].abs().clamp(min=1.0)  # no near-tie
    top_poly = grid.candidates.view(-1, 16)[c, top].to(torch.int64)
    want_poly, want_pt = exact_pairs(env, P, xy)
    per_point = ranker.per_point_polygon(want_poly, want_pt, len(x)).to(torch.int64)
    loc = ranker.LearnedLocator(P, grid, model, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    res = loc.locate(cs.Points.from_points_xy(xy), max_tests=1)
    first_right = (res.flags & 2) == 0
    gemm_right = top_poly == per_point
    sel = clear & (per_point >= 0) & (cnt > 1) & (cnt <= 16)
    agree = (first_right[sel] == gemm_right[sel]).to(torch.float64).mean().item()
    assert sel.sum().item() > 50000 and agree > 0.999, agree


def test_locator_rejects_overlapping_polygons(env):
    """The locator answers with the first confirmed candidate, which is the plain
    join's answer only for non-overlapping polygons: overlapping star polygons
    (and a polygon nested in another) are refused when the locator is built,
    while the Voronoi cells with holes filled by islands are accepted."""
    from conftest import star_polygons

    torch, cs, ranker, rk, datasets = env
    rng = np.random.default_rng(41)
    part, ring, vx, vy = star_polygons(rng, 300, 1e5, 2000.0, 6000.0)
    P = cs.Polygons(part, ring, vx, vy, device="cuda")
    grid = ranker.CandidateGrid.build(P, 0.0, 0.0, 1e5, 32)
    oi, oj = ranker.overlapping_polygons(P, grid)
    assert oi.numel() > 0
    with pytest.raises(ValueError, match="non-overlapping"):
        ranker.LearnedLocator(P, grid, ranker.Ranker(grid.num_cells), (0.0, 1e5, 0.0, 1e5),
                              1e5 / 1024, 10, 64)
    # a square nested inside another, with no edge crossings
    sq = lambda x0, y0, s: ([x0, x0 + s, x0 + s, x0, x0], [y0, y0, y0 + s, y0 + s, y0])  # noqa
    ax, ay = sq(1000.0, 1000.0, 5000.0)
    bx, by = sq(2000.0, 2000.0, 500.0)
    Pn = cs.Polygons(np.array([0, 1, 2], np.int32), np.array([0, 5, 10], np.int32),
                     np.array(ax + bx), np.array(ay + by), device="cuda")
    gn = ranker.CandidateGrid.build(Pn, 0.0, 0.0, 1e5, 16)
    oi, oj = ranker.overlapping_polygons(Pn, gn)
    assert oi.tolist() == [0] and oj.tolist() == [1]
    # a partition passes
    _, Pv, gv, _, _, _ = voronoi(env, 10, 500, 42, 32)
    assert ranker.overlapping_polygons(Pv, gv)[0].numel() == 0


def test_locator_restart_from_file(env, tmp_path):
    """A trained locator saved to one index file (polygons, grid, ranker weights,
    fallback parameters) and loaded in a fresh process answers every point
    exactly as before: identical pairs, tests per point and miss flags."""
    import subprocess
    import sys

    torch, cs, ranker, rk, _ = env
    _, P, grid, xy, x, y = voronoi(env, 100000, 400, 43, 32)
    want_poly, want_pt = exact_pairs(env, P, xy)
    labels = grid.slots_of(xy, ranker.per_point_polygon(want_poly, want_pt, len(x)))
    model = ranker.Ranker(grid.num_cells, seed=44)
    model.train(grid, xy, labels, epochs=3, batch=8192, lr=5e-3)
    loc = ranker.LearnedLocator(P, grid, model, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    res = loc.locate(cs.Points.from_points_xy(xy))
    poly, pt = res.pairs()
    assert torch.equal(poly, want_poly) and torch.equal(pt, want_pt)
    path = tmp_path / "locator.arrow"
    loc.save(path)
    np.save(tmp_path / "xy.npy", xy.cpu().numpy())
    code = (
        "import sys, numpy as np, torch\n"
        f"sys.path.insert(0, {repr(str(__import__('os').path.dirname(__import__('os').path.dirname(__file__))))})\n"
        "import cuspatial_amd as cs\n"
        "from cuspatial_amd import ranker\n"
        f"loc = ranker.LearnedLocator.load({str(path)!r})\n"
        f"xy = torch.as_tensor(np.load({str(tmp_path / 'xy.npy')!r}), device='cuda')\n"
        "r = loc.locate(cs.Points.from_points_xy(xy))\n"
        "poly, pt = r.pairs()\n"
        f"np.savez({str(tmp_path / 'out.npz')!r}, poly=poly.cpu().numpy(), pt=pt.cpu().numpy(), "
        "tests=r.tests.cpu().numpy(), flags=r.flags.cpu().numpy())\n"
        "print('restart ok', int(poly.numel()))\n")
    print("fresh process: load the saved locator and locate 100000 points", flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.load(tmp_path / "out.npz")
    assert np.array_equal(got["poly"], poly.cpu().numpy())
    assert np.array_equal(got["pt"], pt.cpu().numpy())
    assert np.array_equal(got["tests"], res.tests.cpu().numpy())
    assert np.array_equal(got["flags"], res.flags.cpu().numpy())
