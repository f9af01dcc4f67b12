"""The benchmark configurations at their configured shapes, on the GPU.

  C2  1e8 Gaussian-mixture points, 50,000 star polygons with holes, depth 15,
      max_size 256: every output column of the four entry points against the
      oracle (point_indices, the five quadtree columns, the join pairs, and all
      (polygon, point) hits).
  C3  4,096 jobs (5.9e8 points, 1.08M polygons, 4-256 vertices): the batched
      call against the four single-job calls for every job, against the oracle
      for every 64th job, and the captured plan's replays against the eager call.
  C4  2e8 points in the 20,000-cell Voronoi partition, 256 x 256 grid, ranker
      trained on the GPU on 1e7 points labelled by the plain join: the learned
      locator's pairs against the plain join's for all points, and against the
      oracle on every 997th point; the trained model misses less than list order.
  C5  the resident 50,000 star polygons of C2 and the Voronoi partition of their
      centres, requests log-uniform in [1, 1e6] points: every request's answer is
      the same alone and in the mix, and the oracle's for a few of them.

Bar: bit-identical integer columns, as everywhere in this suite.
"""

import gc

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd as cs
    from cuspatial_amd import datasets, ranker, serving

    return torch, cs, datasets, ranker, serving


@pytest.fixture(autouse=True)
def _release_memory():
    yield
    import torch

    from cuspatial_amd import _capi

    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    _capi.lib().qpip_trim()


def u32(t):
    import torch

    return t.view(torch.int32).cpu().numpy().view(np.uint32)


def test_c2_every_column_matches_oracle(env, oracle_mod):
    torch, cs, datasets, _, _ = env
    w = datasets.gaussian_holes(device="cuda")
    print(f"C2: {w.num_points} points, {w.num_polygons} polygons, depth {w.max_depth}",
          flush=True)
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    pts = cs.Points.from_points_xy(torch.stack([w.x, w.y], 1).reshape(-1))
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    pi, tree = cs.quadtree_on_points(pts, *bbox, w.scale, w.max_depth, w.max_size)
    bb = cs.polygon_bounding_boxes(polys)
    pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *bbox, w.scale, w.max_depth)
    res = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, polys)
    torch.cuda.synchronize()
    del pts
    x, y = w.x.cpu().numpy(), w.y.cpu().numpy()
    print("C2: oracle over all points", flush=True)
    o_pi, key, level, internal, length, offset = oracle_mod.quadtree_on_points(
        x, y, *bbox, w.scale, w.max_depth, w.max_size)
    assert np.array_equal(u32(pi), o_pi)
    del o_pi
    for name, want in (("key", key), ("level", level), ("length", length), ("offset", offset)):
        got = tree[name].cpu().numpy()
        assert np.array_equal(got.astype(np.int64), want.astype(np.int64)), name
    assert np.array_equal(tree["is_internal_node"].cpu().numpy().astype(bool), internal)
    otree = (key, level, internal, length, offset)
    obb = oracle_mod.polygon_bounding_boxes(w.part_offset, w.ring_offset, w.px, w.py)
    jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(otree, obb, *bbox, w.scale, w.max_depth)
    assert np.array_equal(u32(pairs["bbox_offset"]), jb)
    assert np.array_equal(u32(pairs["quad_offset"]), jq)
    opoly, opt = oracle_mod.quadtree_point_in_polygon((jb, jq), otree, u32(pi), x, y,
                                                      w.part_offset, w.ring_offset, w.px, w.py)
    print(f"C2: {len(opoly)} hits", flush=True)
    assert len(opoly) > 1_000_000
    assert np.array_equal(u32(res["polygon_index"]), opoly)
    assert np.array_equal(u32(res["point_index"]), opt)


def test_c3_4096_jobs_match_single_calls_and_oracle(env, oracle_mod):
    import torch

    _, cs, datasets, _, _ = env
    w = datasets.batched_jobs(device="cuda")
    assert w.num_jobs == 4096
    pts = cs.Points.from_points_xy(w.xy)
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    print(f"C3: {w.num_jobs} jobs, {w.num_points} points, {len(polys)} polygons", flush=True)
    out = cs.batched_quadtree_point_in_polygon(pts, w.job_point_offsets, polys,
                                               w.job_polygon_offsets, w.bboxes, w.scales,
                                               w.max_depths, w.max_sizes)
    # every job against the four single-job calls on its own points and polygons
    part_h = np.asarray(w.part_offset, np.int64)
    ring_h = np.asarray(w.ring_offset, np.int64)
    part_d = torch.as_tensor(w.part_offset, device="cuda")
    ring_d = torch.as_tensor(w.ring_offset, device="cuda")
    for j in range(w.num_jobs):
        a, b = int(w.job_point_offsets[j]), int(w.job_point_offsets[j + 1])
        q0, q1 = int(w.job_polygon_offsets[j]), int(w.job_polygon_offsets[j + 1])
        r0, r1 = int(part_h[q0]), int(part_h[q1])
        v0, v1 = int(ring_h[r0]), int(ring_h[r1])
        P = cs.Polygons((part_d[q0:q1 + 1] - r0).to(torch.int32),
                        (ring_d[r0:r1 + 1] - v0).to(torch.int32), polys.x[v0:v1], polys.y[v0:v1])
        pj = cs.Points.from_points_xy(w.xy[2 * a:2 * b])
        x0, x1, y0, y1 = w.bboxes[j]
        d, ms, sc = int(w.max_depths[j]), int(w.max_sizes[j]), float(w.scales[j])
        pi, tree = cs.quadtree_on_points(pj, x0, x1, y0, y1, sc, d, ms)
        pairs = cs.join_quadtree_and_bounding_boxes(tree, cs.polygon_bounding_boxes(P), x0, x1,
                                                    y0, y1, sc, d)
        res = cs.quadtree_point_in_polygon(pairs, tree, pi, pj, P)
        bpi, bres = out.job(j)
        assert torch.equal(pi.view(torch.int32), bpi.view(torch.int32)), j
        for k in ("polygon_index", "point_index"):
            assert torch.equal(res[k].view(torch.int32), bres[k].view(torch.int32)), (j, k)
    print("C3: all jobs equal the single-job calls", flush=True)
    # every 64th job against the oracle
    for j in range(0, w.num_jobs, 64):
        x, y = w.job_points(j)
        part, ring, px, py = w.job_polygons(j)
        x0, x1, y0, y1 = w.bboxes[j]
        d, ms, sc = int(w.max_depths[j]), int(w.max_sizes[j]), float(w.scales[j])
        o = oracle_mod.quadtree_on_points(x, y, x0, x1, y0, y1, sc, d, ms)
        obb = oracle_mod.polygon_bounding_boxes(part, ring, px, py)
        jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(o[1:], obb, x0, x1, y0, y1, sc, d)
        opoly, opt = oracle_mod.quadtree_point_in_polygon((jb, jq), o[1:], o[0], x, y, part,
                                                          ring, px, py)
        bpi, bres = out.job(j)
        assert np.array_equal(u32(bpi), o[0]), j
        assert np.array_equal(u32(bres["polygon_index"]), opoly), j
        assert np.array_equal(u32(bres["point_index"]), opt), j
    print("C3: every 64th job equals the oracle", flush=True)
    plan = cs.BatchPlan(pts, w.job_point_offsets, polys, w.job_polygon_offsets, w.bboxes,
                        w.scales, w.max_depths, w.max_sizes)
    for _ in range(3):
        plan.launch()
        torch.cuda.synchronize()
        got = plan.result()
        assert np.array_equal(got.job_hit_offsets, out.job_hit_offsets)
        for k in ("point_indices", "polygon_index", "point_index"):
            assert torch.equal(getattr(got, k).view(torch.int32),
                               getattr(out, k).view(torch.int32)), k
    plan.close()


def plain_pairs_all(env, pts, polys, w):
    _, cs, _, ranker, _ = env
    return ranker.plain_pairs(pts, polys, (w.x_min, w.x_max, w.y_min, w.y_max), w.scale,
                              w.max_depth, w.max_size)


def trained_locator(env, polys, w, xy, n, train_points):
    torch, cs, _, ranker, _ = env
    grid = ranker.CandidateGrid.build(polys, w.x_min, w.y_min, w.x_max - w.x_min, 256)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(w.seed)
    sample = torch.randperm(n, generator=gen, device="cuda")[:train_points]
    txy = xy.view(-1, 2)[sample].reshape(-1).contiguous()
    poly, pt = plain_pairs_all(env, cs.Points.from_points_xy(txy), polys, w)
    labels = grid.slots_of(txy, ranker.per_point_polygon(poly, pt, train_points))
    model = ranker.Ranker(grid.num_cells, seed=1)
    hist = model.train(grid, txy, labels, epochs=3, batch=16384, lr=3e-3)
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    return ranker.LearnedLocator(polys, grid, model, bbox, w.scale, w.max_depth,
                                 w.max_size), grid, hist


def test_c4_learned_locator_at_config_shape(env, oracle_mod):
    torch, cs, datasets, ranker, _ = env
    w = datasets.voronoi_partition(device="cuda")
    n = w.num_points
    assert n == 200_000_000 and w.num_polygons == 20_000
    xy = torch.stack([w.x, w.y], 1).reshape(-1)
    pts = cs.Points.from_points_xy(xy)
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    print(f"C4: {n} points, {w.num_polygons} cells; training on 1e7 points", flush=True)
    loc, grid, hist = trained_locator(env, polys, w, xy, n, 10_000_000)
    assert hist[-1][1] > hist[0][1], hist
    want_poly, want_pt = plain_pairs_all(env, pts, polys, w)
    res = loc.locate(pts)
    poly, pt = res.pairs()
    assert torch.equal(poly, want_poly) and torch.equal(pt, want_pt)
    per_point = ranker.per_point_polygon(want_poly, want_pt, n)
    has = per_point >= 0
    miss = float((((res.flags & 2) > 0) & has).sum()) / float(has.sum())
    zero = ranker.LearnedLocator(polys, grid, ranker.Ranker(grid.num_cells, zero=True),
                                 loc.bbox, w.scale, w.max_depth, w.max_size,
                                 check_partition=False)
    zres = zero.locate(pts)
    zmiss = float((((zres.flags & 2) > 0) & has).sum()) / float(has.sum())
    print(f"C4: miss rate {miss:.4f} (list order {zmiss:.4f}); tests/point "
          f"{res.tests.to(torch.float64).mean().item():.3f}", flush=True)
    assert miss < 0.5 * zmiss
    # the oracle on every 997th point, with all 20,000 cells
    idx = torch.arange(0, n, 997, device="cuda")
    sx, sy = w.x[idx].cpu().numpy(), w.y[idx].cpu().numpy()
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    o = oracle_mod.quadtree_on_points(sx, sy, *bbox, w.scale, w.max_depth, w.max_size)
    obb = oracle_mod.polygon_bounding_boxes(w.part_offset, w.ring_offset, w.px, w.py)
    jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(o[1:], obb, *bbox, w.scale, w.max_depth)
    opoly, opos = oracle_mod.quadtree_point_in_polygon((jb, jq), o[1:], o[0], sx, sy,
                                                       w.part_offset, w.ring_offset, w.px, w.py)
    want = np.full(len(sx), -1, np.int64)
    want[o[0][opos].astype(np.int64)] = opoly.astype(np.int64)
    got = res.polygon[idx].cpu().numpy().astype(np.int64)
    assert np.array_equal(got, want)
    # the partition covers the domain: only points on a shared edge have no cell
    assert int((want >= 0).sum()) >= len(sx) - 2


def test_c5_serving_at_config_shape(env, oracle_mod):
    torch, cs, datasets, ranker, serving = env
    gh = datasets.gaussian_holes(num_points=1, device="cuda")
    bbox = (gh.x_min, gh.x_max, gh.y_min, gh.y_max)
    stars = cs.Polygons(gh.part_offset, gh.ring_offset, gh.px, gh.py)
    plain_index = serving.ResidentIndex(stars, bbox, gh.scale, gh.max_depth, gh.max_size)
    vw = datasets.gaussian_holes_partition(num_points=2_000_000, device="cuda")
    vor = cs.Polygons(vw.part_offset, vw.ring_offset, vw.px, vw.py)
    vxy = torch.stack([vw.x, vw.y], 1).reshape(-1)
    loc, _, _ = trained_locator(env, vor, vw, vxy, vxy.numel() // 2, 2_000_000)
    learned_index = serving.ResidentIndex(vor, bbox, vw.scale, vw.max_depth, vw.max_size,
                                          locator=loc)
    px, py = datasets._gaussian_mixture_device(2005, 20_000_000, 1e5, "cuda")
    pool = torch.stack([px, py], 1).reshape(-1)
    del px, py
    reqs_of = lambda: serving.poisson_requests(pool, qps=20000, count=48, seed=5,  # noqa: E731
                                               max_points=1_000_000)
    for name, index, (part, ring, vx, vy) in (
            ("plain_stars", plain_index, (gh.part_offset, gh.ring_offset, gh.px, gh.py)),
            ("learned_voronoi", learned_index, (vw.part_offset, vw.ring_offset, vw.px, vw.py))):
        reqs = reqs_of()
        assert max(r.points for r in reqs) > 100_000
        st = serving.Server(index, workers=3, keep_results=True).serve(reqs)
        print(f"C5 {name}: {st.requests} requests, {st.points} points, {st.batches} batches, "
              f"p99 {st.latency_ms['p99']:.2f} ms", flush=True)
        assert st.requests == len(reqs) and st.mean_batch_requests > 1.0
        for r in reqs:
            p, t = index.query(r.xy)
            assert torch.equal(p, r.result[0]) and torch.equal(t, r.result[1]), (name, r.rid)
        small = sorted(reqs, key=lambda r: r.points)[len(reqs) // 2:len(reqs) // 2 + 3]
        for r in small:
            a = r.xy.view(-1, 2).cpu().numpy()
            x, y = a[:, 0].copy(), a[:, 1].copy()
            o = oracle_mod.quadtree_on_points(x, y, *bbox, gh.scale, gh.max_depth, gh.max_size)
            obb = oracle_mod.polygon_bounding_boxes(part, ring, vx, vy)
            jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(o[1:], obb, *bbox, gh.scale,
                                                                 gh.max_depth)
            opoly, opos = oracle_mod.quadtree_point_in_polygon((jb, jq), o[1:], o[0], x, y, part,
                                                               ring, vx, vy)
            opt = o[0][opos].astype(np.int64)
            order = np.lexsort((opoly.astype(np.int64), opt))
            assert np.array_equal(r.result[1].cpu().numpy(), opt[order]), (name, r.rid)
            assert np.array_equal(r.result[0].cpu().numpy(), opoly.astype(np.int64)[order])
