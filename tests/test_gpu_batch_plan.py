"""The batched join captured as one HIP graph (BatchPlan) against the eager call.

Bar: every replay's point_indices, (polygon_index, point_index) hits and per-job
hit offsets are identical to batched_quadtree_point_in_polygon on the same points
(which test_gpu_batch.py checks against the oracle job by job). Replays run
several times with eager work of the program in between, after the point buffer
is rewritten in place, and across several sort groups; a replay whose points need
more room than the plan holds reports CapacityError and leaves later calls
correct.
"""

import numpy as np
import pytest

from test_gpu_batch import concat_jobs, star_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cs():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd

    return cuspatial_amd


def device_inputs(cs, w):
    import torch

    xy = torch.as_tensor(np.asarray(w.xy, np.float64), device="cuda").contiguous()
    P = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py, device="cuda")
    return xy, P


def eager(cs, w, xy, P):
    return cs.batched_quadtree_point_in_polygon(
        cs.Points.from_points_xy(xy), w.job_point_offsets, P, w.job_polygon_offsets, w.bboxes,
        w.scales, w.max_depths, w.max_sizes)


def make_plan(cs, w, xy, P, headroom=1.25):
    return cs.BatchPlan(cs.Points.from_points_xy(xy), w.job_point_offsets, P,
                        w.job_polygon_offsets, w.bboxes, w.scales, w.max_depths, w.max_sizes,
                        headroom=headroom)


def assert_same(a, b):
    import torch

    assert np.array_equal(a.job_hit_offsets, b.job_hit_offsets)
    for k in ("point_indices", "polygon_index", "point_index"):
        x, y = getattr(a, k), getattr(b, k)
        assert x.numel() == y.numel(), k
        assert torch.equal(x.view(torch.int32).cpu(), y.view(torch.int32).cpu()), k


def new_points(w, seed):
    """Fresh points, uniform in each job's bbox, with the same counts per job."""
    rng = np.random.default_rng(seed)
    out = np.empty(2 * int(w.job_point_offsets[-1]), np.float64)
    for j in range(w.num_jobs):
        a, b = int(w.job_point_offsets[j]), int(w.job_point_offsets[j + 1])
        x0, x1, y0, y1 = w.bboxes[j]
        out[2 * a:2 * b:2] = rng.uniform(min(x0, x1), max(x0, x1), b - a)
        out[2 * a + 1:2 * b:2] = rng.uniform(min(y0, y1), max(y0, y1), b - a)
    return out


def test_plan_replays_match_eager(cs):
    import torch

    from cuspatial_amd import datasets

    w = datasets.batched_jobs(seed=21, num_jobs=96, min_points=10, max_points=60000,
                              min_polygons=1, max_polygons=200, max_vertices=64, ext=1e4)
    xy, P = device_inputs(cs, w)
    print(f"[plan] jobs={w.num_jobs} points={w.num_points}", flush=True)
    ref = eager(cs, w, xy, P)
    assert ref.polygon_index.numel() > 0
    plan = make_plan(cs, w, xy, P)
    for r in range(3):
        plan.launch()
        got = plan.result()
        assert_same(got, ref)
        # the program's own eager work between replays
        again = eager(cs, w, xy, P)
        assert_same(again, ref)
        torch.cuda.synchronize()
        print(f"[plan] replay {r} ok: {got.polygon_index.numel()} hits", flush=True)
    plan.close()


def test_plan_rewritten_points_and_groups(cs):
    import torch

    rng = np.random.default_rng(31)
    jobs = [star_job(rng, 5000, 30), star_job(rng, 0, 5),
            dict(star_job(rng, 3000, 0), part=[0], ring=[0], vx=[], vy=[]),
            star_job(rng, 20000, 20, d=15, max_size=5),  # a sort group of its own
            star_job(rng, 1, 1, d=1, max_size=1), star_job(rng, 60000, 40, d=9, max_size=64),
            star_job(rng, 30000, 25, d=15, max_size=40)]
    w = concat_jobs(jobs)
    xy, P = device_inputs(cs, w)
    plan = make_plan(cs, w, xy, P, headroom=2.0)
    plan.launch()
    assert_same(plan.result(), eager(cs, w, xy, P))
    for seed in (1, 2):
        xy.copy_(torch.as_tensor(new_points(w, seed), device="cuda"))
        plan.launch()
        got = plan.result()
        ref = eager(cs, w, xy, P)
        assert ref.polygon_index.numel() > 0
        assert_same(got, ref)
        print(f"[plan] rewritten points {seed}: {got.polygon_index.numel()} hits", flush=True)
    plan.close()


def test_plan_capacity_exceeded_is_reported(cs):
    import torch

    rng = np.random.default_rng(41)
    job = star_job(rng, 50000, 60, d=7, max_size=32)
    # sized on points packed into one corner: few leaves meet polygons
    job["x"] = rng.uniform(0.0, 0.5, 50000)
    job["y"] = rng.uniform(0.0, 0.5, 50000)
    w = concat_jobs([job])
    xy, P = device_inputs(cs, w)
    plan = make_plan(cs, w, xy, P, headroom=1.0)
    plan.launch()
    assert_same(plan.result(), eager(cs, w, xy, P))
    spread = new_points(w, 5)
    xy.copy_(torch.as_tensor(spread, device="cuda"))
    ref = eager(cs, w, xy, P)
    assert ref.polygon_index.numel() > 5000  # far past the plan's hit capacity
    plan.launch()
    with pytest.raises(cs.CapacityError, match="capacity exceeded"):
        plan.result()
    # nothing else was disturbed: eager calls and a new plan give the same answer
    assert_same(eager(cs, w, xy, P), ref)
    plan.close()
    plan2 = make_plan(cs, w, xy, P)
    plan2.launch()
    assert_same(plan2.result(), ref)
    plan2.close()


def test_plan_under_each_saved_tuning_matches_eager(cs, tmp_path):
    """A plan looks its shape bucket up in a tuning table and is captured under
    those values; every configuration of the table (and of the package's own
    table, if it has any) gives the eager call's results exactly, and the
    process-wide values are restored after the plan is made."""
    from cuspatial_amd import autotune, datasets

    w = datasets.batched_jobs(seed=23, num_jobs=40, min_points=100, max_points=40000,
                              min_polygons=1, max_polygons=150, max_vertices=64, ext=1e4)
    xy, P = device_inputs(cs, w)
    want = eager(cs, w, xy, P)
    configs = [
        {"pip_min_pairs_per_wave": 1, "pip_pairs_shift": 10, "pip_waves_per_block": 8,
         "pip_stage_edges": 128, "pip_xcd_run": 1},
        {"pip_min_pairs_per_wave": 64, "pip_pairs_shift": 16, "pip_waves_per_block": 4,
         "pip_stage_edges": 64, "pip_xcd_run": 4},
        {"pip_min_pairs_per_wave": 1, "pip_pairs_shift": 24, "pip_waves_per_block": 2,
         "pip_stage_edges": 128, "pip_polygon_group": 16},
        {"pip_min_pairs_per_wave": 1, "pip_pairs_shift": 24, "pip_waves_per_block": 8,
         "pip_stage_edges": 64, "pip_polygon_group": 2, "pip_box_filter": 0},
    ]
    if __import__("os").path.exists(autotune.DEFAULT_TABLE):
        configs += [dict(b["config"]) for b in autotune.TuningTable(autotune.DEFAULT_TABLE)
                    .buckets.values()]
    before = autotune.current()
    for k, cfg in enumerate(configs):
        table = autotune.TuningTable(None)
        probe = make_plan(cs, w, xy, P)
        bucket = probe.bucket
        probe.close()
        table.record(bucket, {"config": cfg, "baseline_ms": 0.0, "best_ms": 0.0}, "test")
        path = str(tmp_path / f"t{k}.json")
        table.save(path)
        print(f"[plan] tuning {cfg} for bucket {bucket}", flush=True)
        plan = cs.BatchPlan(cs.Points.from_points_xy(xy), w.job_point_offsets, P,
                            w.job_polygon_offsets, w.bboxes, w.scales, w.max_depths,
                            w.max_sizes, tuning=path)
        assert plan.tuning == cfg
        assert autotune.current() == before
        plan.launch()
        plan.launch()
        assert_same(plan.result(), want)
        plan.close()
