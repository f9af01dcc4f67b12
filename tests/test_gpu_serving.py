"""The serving loop: every request's answer is the same alone and in the mix
(exact: the (polygon, point) pairs are bit-identical), batches coalesce and
overlap on the device, and the memory budget holds."""

import numpy as np
import pytest

from conftest import star_polygons

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd as cs
    from cuspatial_amd import datasets, ranker, serving

    return torch, cs, ranker, serving, datasets


def pool(torch, n, seed, ext):
    rng = np.random.default_rng(seed)
    return torch.as_tensor(rng.uniform(0, ext, 2 * n), device="cuda")


def star_index(env):
    torch, cs, ranker, serving, _ = env
    rng = np.random.default_rng(41)
    part, ring, vx, vy = star_polygons(rng, 300, 1000.0, 5.0, 60.0, holes=0.3)
    P = cs.Polygons(part, ring, vx, vy, device="cuda")
    return serving.ResidentIndex(P, (0.0, 1000.0, 0.0, 1000.0), 1000.0 / 256, 8, 64), \
        (part, ring, vx, vy)


def check_alone_equals_mix(env, index, reqs):
    torch = env[0]
    for r in reqs:
        poly, pt = index.query(r.xy)
        assert torch.equal(poly, r.result[0]) and torch.equal(pt, r.result[1]), r.rid


def test_star_index_alone_equals_mix(env, oracle_mod):
    torch, cs, ranker, serving, _ = env
    index, (part, ring, vx, vy) = star_index(env)
    xy = pool(torch, 400000, 42, 1000.0)
    reqs = serving.poisson_requests(xy, qps=50000, count=80, seed=1, max_points=20000)
    srv = serving.Server(index, workers=3, keep_results=True)
    st = srv.serve(reqs)
    assert st.requests == 80 and all(r.result is not None for r in reqs)
    assert st.mean_batch_requests > 1.0, st  # coalesced
    assert st.max_in_flight >= 2, st  # overlapped on the device
    check_alone_equals_mix(env, index, reqs)
    # and the answer is the oracle's join for a few requests
    for r in reqs[:3]:
        a = r.xy.view(-1, 2).cpu().numpy()
        x, y = a[:, 0].copy(), a[:, 1].copy()
        pi, key, level, internal, length, offset = oracle_mod.quadtree_on_points(
            x, y, 0.0, 1000.0, 0.0, 1000.0, 1000.0 / 256, 8, 64)
        tree = (key, level, internal, length, offset)
        bb = oracle_mod.polygon_bounding_boxes(part, ring, vx, vy)
        jb, jq = oracle_mod.join_quadtree_and_bounding_boxes(tree, bb, 0.0, 1000.0, 0.0, 1000.0,
                                                             1000.0 / 256, 8)
        pp, pt = oracle_mod.quadtree_point_in_polygon((jb, jq), tree, pi, x, y, part, ring, vx, vy)
        opt = np.asarray(pi, np.int64)[np.asarray(pt, np.int64)]
        order = np.lexsort((np.asarray(pp, np.int64), opt))
        assert r.result[0].cpu().numpy().tolist() == np.asarray(pp, np.int64)[order].tolist()
        assert r.result[1].cpu().numpy().tolist() == opt[order].tolist()


def test_learned_index_alone_equals_mix(env):
    torch, cs, ranker, serving, datasets = env
    part, ring, px, py = datasets.voronoi_cells(43, 400, 1e5)
    P = cs.Polygons(part, ring, px, py, device="cuda")
    grid = ranker.CandidateGrid.build(P, 0.0, 0.0, 1e5, 32)
    loc = ranker.LearnedLocator(P, grid, ranker.Ranker(grid.num_cells, seed=44),
                                (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    index = serving.ResidentIndex(P, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64, locator=loc)
    plain = serving.ResidentIndex(P, (0.0, 1e5, 0.0, 1e5), 1e5 / 1024, 10, 64)
    xy = pool(torch, 400000, 45, 1e5)
    reqs = serving.poisson_requests(xy, qps=50000, count=60, seed=2, max_points=20000)
    st = serving.Server(index, workers=3, keep_results=True).serve(reqs)
    assert st.requests == 60 and st.mean_batch_requests > 1.0
    check_alone_equals_mix(env, index, reqs)
    check_alone_equals_mix(env, plain, reqs)  # and the plain join gives the same pairs


def test_memory_budget_admission(env):
    """With a budget smaller than most requests, big requests run alone and the
    in-flight bytes never exceed the budget except for a lone oversize request."""
    torch, cs, ranker, serving, _ = env
    index, _ = star_index(env)
    xy = pool(torch, 200000, 46, 1000.0)
    reqs = serving.poisson_requests(xy, qps=100000, count=40, seed=3, max_points=50000)
    srv = serving.Server(index, workers=3, memory_budget=256 * 5000, bytes_per_point=256,
                         keep_results=True)
    st = srv.serve(reqs)
    assert st.requests == 40
    biggest = max(r.points for r in reqs) * 256
    assert st.extra["max_in_flight_bytes"] <= max(256 * 5000, biggest)
    check_alone_equals_mix(env, index, reqs)
    assert np.all(np.isfinite([r.done for r in reqs]))


def test_concurrent_calls_match_serial(env):
    """Several host threads calling the library at once, each on its own stream:
    every result equals the same call made alone. (The library's temporaries are
    reused only in the stream order of the thread that freed them.)"""
    import threading

    torch, cs, ranker, serving, _ = env
    index, _ = star_index(env)
    xys = [pool(torch, 1_500_000, 50 + k, 1000.0) for k in range(4)]
    want = [index.query(xy) for xy in xys]
    torch.cuda.synchronize()
    got = [[None] * 3 for _ in xys]
    errors = []

    def run(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for it in range(3):
                    got[k][it] = index.query(xys[k])
                s.synchronize()
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    th = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k in range(4):
        for it in range(3):
            assert torch.equal(got[k][it][0], want[k][0]) and torch.equal(got[k][it][1], want[k][1])


def test_captured_batches_equal_eager_batches(env):
    """The plain index answers a batch by replaying a captured plan of the batch's
    size bucket (a point outside the bounding box fills the bucket): the pairs
    equal the eager path's
    for batch sizes across bucket edges, and a batch that outgrows the plan's
    capacities (many more hits than the batch the plan was sized on) rebuilds the
    plan and still answers exactly."""
    torch, cs, ranker, serving, _ = env
    captured, (part, ring, vx, vy) = star_index(env)
    P = captured.polygons
    eager = serving.ResidentIndex(P, captured.bbox, captured.scale, captured.max_depth,
                                  captured.max_size, captured=False)
    rng = np.random.default_rng(47)
    for n in (1, 100, 4095, 4096, 4097, 20000, 70000):
        xy = torch.as_tensor(rng.uniform(0, 1000.0, 2 * n), device="cuda")
        a, b = captured.query(xy), eager.query(xy)
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), n
    assert captured.plan_stats["replays"] == 7 and eager.plan_stats["eager"] == 7
    # plans created ahead of serving: every bucket up to 2^14 on two streams,
    # then queries of those sizes replay them without a capture
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    pool = torch.as_tensor(rng.uniform(0, 1000.0, 2 * 20000), device="cuda")
    created = captured.plan_stats["created"]
    captured.warm(pool, streams, max_points=1 << 14)
    assert captured.plan_stats["created"] == created + 2 * 3  # 2^12, 2^13, 2^14
    for s in streams:
        with torch.cuda.stream(s):
            a = captured.query(pool[:2 * 9000])
        s.synchronize()
        b = eager.query(pool[:2 * 9000])
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    assert captured.plan_stats["created"] == created + 6
    # same bucket, sized on uniform points, then every point inside one polygon
    xy = torch.as_tensor(rng.uniform(0, 1000.0, 2 * 5000), device="cuda")
    captured.query(xy)
    i = int(np.argmax(np.diff(ring)))  # a ring with many vertices
    cx, cy = float(np.mean(vx[ring[i]:ring[i + 1] - 1])), float(np.mean(vy[ring[i]:ring[i + 1] - 1]))
    dense = np.stack([cx + rng.uniform(-0.5, 0.5, 5000), cy + rng.uniform(-0.5, 0.5, 5000)], 1)
    xy = torch.as_tensor(dense.reshape(-1), device="cuda")
    a, b = captured.query(xy), eager.query(xy)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    assert b[0].numel() > 0
    captured.close()
