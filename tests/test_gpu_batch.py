"""Parity of the batched join (many jobs in one pass) with the oracle run job by job.

Bar: for every job, point_indices and the (polygon_index, point_index) hits are
bit-identical to the oracle's single-job pipeline (quadtree_on_points ->
polygon_bounding_boxes -> join -> quadtree_point_in_polygon) on the job's own
points and polygons.
"""

import numpy as np
import pytest

from conftest import star_polygons

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cs():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd

    return cuspatial_amd


def np_(t):
    return t.cpu().numpy()


def concat_jobs(jobs):
    """jobs: dicts with x, y, part, ring, vx, vy, bbox (x_min, x_max, y_min, y_max),
    scale, d, max_size -> BatchWorkload."""
    from cuspatial_amd.datasets import BatchWorkload

    xy, parts, rings, vxs, vys = [], [0], [0], [], []
    jpo, jpy = [0], [0]
    for jb in jobs:
        x, y = np.asarray(jb["x"], np.float64), np.asarray(jb["y"], np.float64)
        xy.append(np.stack([x, y], 1).reshape(-1))
        jpo.append(jpo[-1] + len(x))
        part, ring = np.asarray(jb["part"], np.int64), np.asarray(jb["ring"], np.int64)
        r_base, v_base = len(rings) - 1, len(np.concatenate(vxs)) if vxs else 0
        parts += list(part[1:] + r_base)
        rings += list(ring[1:] + v_base)
        vxs.append(np.asarray(jb["vx"], np.float64))
        vys.append(np.asarray(jb["vy"], np.float64))
        jpy.append(jpy[-1] + len(part) - 1)
    return BatchWorkload(
        "test", np.concatenate(xy) if xy else np.zeros(0), np.array(jpo, np.int64),
        np.array(parts, np.int32), np.array(rings, np.int32),
        np.concatenate(vxs) if vxs else np.zeros(0), np.concatenate(vys) if vys else np.zeros(0),
        np.array(jpy, np.int64), np.array([jb["bbox"] for jb in jobs], np.float64),
        np.array([jb["scale"] for jb in jobs], np.float64),
        np.array([jb["d"] for jb in jobs], np.int8),
        np.array([jb["max_size"] for jb in jobs], np.int32), 0)


def run_batch(cs, w, device="cuda"):
    import torch

    xy = torch.as_tensor(np.asarray(w.xy, np.float64), device=device)
    P = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py, device=device)
    return cs.batched_quadtree_point_in_polygon(
        cs.Points.from_points_xy(xy), w.job_point_offsets, P, w.job_polygon_offsets, w.bboxes,
        w.scales, w.max_depths, w.max_sizes)


def oracle_job(oracle, w, j):
    x, y = w.job_points(j)
    part, ring, vx, vy = w.job_polygons(j)
    x0, x1, y0, y1 = w.bboxes[j]
    d, ms, sc = int(w.max_depths[j]), int(w.max_sizes[j]), float(w.scales[j])
    pi, key, level, internal, length, offset = oracle.quadtree_on_points(
        x, y, x0, x1, y0, y1, sc, d, ms)
    if len(key) == 0 or len(part) <= 1:
        return pi, np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    bb = oracle.polygon_bounding_boxes(part, ring, vx, vy)
    lo_x, hi_x, lo_y, hi_y = min(x0, x1), max(x0, x1), min(y0, y1), max(y0, y1)
    s = max(sc, max(hi_x - lo_x, hi_y - lo_y) / ((1 << d) + 2))
    jb, jq = oracle.join_quadtree_and_bounding_boxes((key, level, internal, length, offset), bb,
                                                     lo_x, hi_x, lo_y, hi_y, s, d)
    pp, pt = oracle.quadtree_point_in_polygon((jb, jq), (key, level, internal, length, offset),
                                              pi, x, y, part, ring, vx, vy)
    return pi, pp, pt


def check_all(cs, oracle, w):
    res = run_batch(cs, w)
    assert len(res) == w.num_jobs
    total = 0
    for j in range(w.num_jobs):
        pi, hits = res.job(j)
        opi, opp, opt = oracle_job(oracle, w, j)
        assert np.array_equal(np_(pi).astype(np.int64), np.asarray(opi).astype(np.int64)), j
        assert np.array_equal(np_(hits["polygon_index"]).astype(np.int64),
                              np.asarray(opp).astype(np.int64)), j
        assert np.array_equal(np_(hits["point_index"]).astype(np.int64),
                              np.asarray(opt).astype(np.int64)), j
        total += len(opp)
    return total


def star_job(rng, n, npoly, ext=100.0, d=6, max_size=20, holes=0.2):
    part, ring, vx, vy = star_polygons(rng, npoly, ext, ext * 0.01, ext * 0.1, 3, 30, holes=holes)
    return dict(x=rng.uniform(0, ext, n), y=rng.uniform(0, ext, n), part=part, ring=ring, vx=vx,
                vy=vy, bbox=(0.0, ext, 0.0, ext), scale=ext / (1 << d), d=d, max_size=max_size)


def test_batch_generated_workload_parity(cs, oracle_mod):
    from cuspatial_amd import datasets

    w = datasets.batched_jobs(seed=7, num_jobs=24, min_points=1, max_points=40000,
                              min_polygons=1, max_polygons=60, max_vertices=40, ext=1e4)
    print(f"[batch] jobs={w.num_jobs} points={w.num_points}", flush=True)
    assert check_all(cs, oracle_mod, w) > 0


def test_batch_edge_cases_parity(cs, oracle_mod):
    rng = np.random.default_rng(11)
    jobs = [star_job(rng, 5000, 30)]
    # no points
    jobs.append(dict(star_job(rng, 0, 5)))
    # no polygons
    jobs.append(dict(star_job(rng, 3000, 0), part=[0], ring=[0], vx=[], vy=[]))
    # an empty polygon (no vertices) in the middle: later boxes shift within the job
    e = star_job(rng, 4000, 6)
    part = list(e["part"][:4]) + [e["part"][3] + 0] + [p + 1 for p in e["part"][4:]]
    ring = list(e["ring"][:e["part"][3] + 1]) + [e["ring"][e["part"][3]]] + \
        list(e["ring"][e["part"][3] + 1:])
    jobs.append(dict(e, part=part, ring=ring))
    # depth 15 (a job whose keys need all 32 bits runs in its own sort group)
    jobs.append(star_job(rng, 20000, 20, d=15, max_size=5))
    # clamped scale and points on the bbox maximum: the root level has several nodes
    q = star_job(rng, 3000, 15, ext=2000.0, d=3, max_size=2)
    q["x"][:40] = 2000.0
    q["y"][20:60] = 2000.0
    q["scale"] = -1.0
    q["bbox"] = (2000.0, 0.0, 0.0, 2000.0)  # unordered
    jobs.append(q)
    # tiny jobs
    jobs += [star_job(rng, 1, 1, d=1, max_size=1), star_job(rng, 2, 3, d=2, max_size=1)]
    jobs.append(star_job(rng, 60000, 40, d=9, max_size=64))
    w = concat_jobs(jobs)
    assert check_all(cs, oracle_mod, w) > 0


@pytest.mark.parametrize("group", [1, 16, 64])
def test_batch_polygon_major_parity(cs, oracle_mod, group):
    """The batched call with the PIP test in polygon-major order (pip_polygon_group)
    and the box filter on, on jobs of several depths (one of depth 15, in a sort
    group of its own) with points on the quarter-cell lines of each job's deepest
    cells (a few ULPs either side) and on the polygons' vertices: the oracle's
    pairs."""
    from cuspatial_amd import autotune

    rng = np.random.default_rng(12)
    jobs = [star_job(rng, 6000, 30), star_job(rng, 20000, 20, d=15, max_size=5),
            star_job(rng, 3000, 15, ext=2000.0, d=3, max_size=2),
            star_job(rng, 60000, 40, d=9, max_size=64)]
    for jb in jobs:
        n, d = len(jb["x"]), jb["d"]
        x0, x1, y0, y1 = jb["bbox"]
        sc = jb["scale"]
        k = n // 4
        ex = x0 + np.floor(rng.uniform(0, 4 << d, k)) * 0.25 * sc
        ey = y0 + np.floor(rng.uniform(0, 4 << d, k)) * 0.25 * sc
        jb["x"][:k] = np.nextafter(ex, np.where(rng.uniform(size=k) < 0.5, -np.inf, np.inf))
        jb["y"][:k] = np.nextafter(ey, np.where(rng.uniform(size=k) < 0.5, -np.inf, np.inf))
        vx, vy = np.asarray(jb["vx"]), np.asarray(jb["vy"])
        if len(vx):
            m = min(n - k, 2 * len(vx))
            i = rng.integers(0, len(vx), m)
            jb["x"][k:k + m] = np.where(rng.uniform(size=m) < 0.5, vx[i], jb["x"][k:k + m])
            jb["y"][k:k + m] = np.where(rng.uniform(size=m) < 0.5, vy[i], jb["y"][k:k + m])
        jb["x"] = np.clip(jb["x"], min(x0, x1), max(x0, x1))
        jb["y"] = np.clip(jb["y"], min(y0, y1), max(y0, y1))
    w = concat_jobs(jobs)
    with autotune.applied({"pip_polygon_group": group, "pip_box_filter": 1}):
        assert check_all(cs, oracle_mod, w) > 0


def test_batch_matches_single_job_calls(cs):
    """The batched result equals the four single-job GPU calls, job by job."""
    import torch

    from cuspatial_amd import datasets

    w = datasets.batched_jobs(seed=9, num_jobs=40, min_points=100, max_points=200000,
                              min_polygons=1, max_polygons=300, max_vertices=64, ext=1e4)
    res = run_batch(cs, w)
    for j in range(w.num_jobs):
        x, y = w.job_points(j)
        part, ring, vx, vy = w.job_polygons(j)
        x0, x1, y0, y1 = w.bboxes[j]
        pts = cs.Points(x, y, device="cuda")
        pi, tree = cs.quadtree_on_points(pts, x0, x1, y0, y1, w.scales[j], int(w.max_depths[j]),
                                         int(w.max_sizes[j]))
        P = cs.Polygons(part, ring, vx, vy, device="cuda")
        bb = cs.polygon_bounding_boxes(P)
        pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, x0, x1, y0, y1, w.scales[j],
                                                    int(w.max_depths[j]))
        hits = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, P)
        bpi, bh = res.job(j)
        assert torch.equal(bpi.cpu().to(torch.int64), pi.cpu().to(torch.int64)), j
        assert torch.equal(bh["polygon_index"].cpu().to(torch.int64),
                           hits["polygon_index"].cpu().to(torch.int64)), j
        assert torch.equal(bh["point_index"].cpu().to(torch.int64),
                           hits["point_index"].cpu().to(torch.int64)), j


def test_batch_errors(cs):
    rng = np.random.default_rng(13)
    w = concat_jobs([star_job(rng, 100, 3), star_job(rng, 100, 3, d=0)])
    with pytest.raises(RuntimeError, match="maximum depth must be positive and less than 16"):
        run_batch(cs, w)
    w = concat_jobs([star_job(rng, 100, 3)])
    w.job_point_offsets[-1] += 1
    with pytest.raises(RuntimeError, match="job point offsets"):
        run_batch(cs, w)
