"""Benchmark of the plain quadtree point-in-polygon join on one MI355X.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]

A step is one pass of the whole path over the gaussian_holes workload with the
inputs resident in HBM: quadtree_on_points -> polygon_bounding_boxes ->
join_quadtree_and_bounding_boxes -> quadtree_point_in_polygon, every stage on
the GPU through libcuspatial_amd's C ABI (outputs included). The metric is
points joined per second. One JSON line is printed; the roofline of the
dominant kernel comes from the library's HIP-event timers over the timed steps,
and the CPU baseline is the oracle (OpenMP) on a bounded prefix of the same
points with the same polygons.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("points/sec joined on 1 MI355X, learned+batched vs plain quadtree PIP (exact); "
          "p99 ms")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
# benchmark (scaled by the same counters on a device copy of known size)
TRAFFIC_FILE = os.path.join(HERE, "profiles", "gaussian_holes_hbm_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def effective_edges(part_offset, ring_offset, px, py):
    """Edges the crossing test visits per polygon: each ring's vertex pairs
    (last->first included) minus zero-length ones (closed rings repeat a vertex)."""
    import numpy as np

    n = px.size
    nxt = np.arange(n) - 1  # b = previous vertex, a = this vertex
    starts = ring_offset[:-1].astype(np.int64)
    ends = ring_offset[1:].astype(np.int64)
    nxt[starts] = ends - 1  # the first vertex pairs with the ring's last
    zero = (px == px[nxt]) & (py == py[nxt])
    ring_id = np.repeat(np.arange(starts.size), ends - starts)
    ring_edges = np.bincount(ring_id, weights=~zero, minlength=starts.size)
    poly_of_ring = np.repeat(np.arange(part_offset.size - 1), np.diff(part_offset))
    return np.bincount(poly_of_ring, weights=ring_edges, minlength=part_offset.size - 1)


def measured_traffic(timer: str, points: int):
    """HBM bytes per launch of the kernel behind `timer` at this point count, from
    the committed counter summary (rows: timer points kernel calls fetch_MiB
    write_MiB), or None when it holds no such row."""
    try:
        with open(TRAFFIC_FILE) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    for ln in lines:
        fields = ln.split()
        if not fields or ln.startswith("#") or len(fields) != 6:
            continue
        if fields[0] == timer and fields[1] == str(points):
            return (float(fields[4]) + float(fields[5])) * 2**20
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=100_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU work of the oracle baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="gaussian_holes",
                    choices=["gaussian_holes", "batched_jobs", "voronoi_learned", "serving"],
                    help="gaussian_holes: the plain join at 1e8 points (the headline line); "
                         "batched_jobs: 4096 independent jobs through the batched entry point, "
                         "with the plain per-job calls on the same inputs timed beside it; "
                         "voronoi_learned: 2e8 points located in a 20,000-cell Voronoi "
                         "partition by the learned ranker + exact test, against the plain "
                         "join on the same inputs; serving: the request loop over both "
                         "resident indexes at a sweep of Poisson request rates")
    ap.add_argument("--qps", default="1000,2000,5000,10000,20000,50000",
                    help="request rates of the serving sweep")
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="arrival window per rate of the serving sweep")
    ap.add_argument("--train-points", type=int, default=10_000_000,
                    help="labelled points the ranker is trained on (voronoi_learned)")
    ap.add_argument("--epochs", type=int, default=3, help="ranker training epochs")
    ap.add_argument("--jobs", type=int, default=4096, help="jobs of the batched workload")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched_jobs: time only the eager batched call, not the captured plan")
    ap.add_argument("--graph-only", action="store_true",
                    help="batched_jobs: one eager call (the reference), then only the captured "
                         "plan's replays (no per-job baseline)")
    ap.add_argument("--autotune", action="store_true",
                    help="batched_jobs: pick the PIP segment parameters with the autotuner "
                         "(cuspatial_amd/autotune.py) on this workload before timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (the known "
                         "traffic that scales the HBM counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path has no "
              "multi-GPU mode in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import _capi, datasets

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calibration} MiB three times")
        return
    if args.workload == "batched_jobs":
        return bench_batched(args, dev)
    if args.workload == "voronoi_learned":
        return bench_learned(args, dev)
    if args.workload == "serving":
        return bench_serving(args, dev)
    log(f"[bench] generating gaussian_holes with {args.points} points on {dev}")
    w = datasets.gaussian_holes(num_points=args.points, device=dev)
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    # points as a GeoSeries holds them: one interleaved x, y array (vec_2d)
    pts = cs.Points.from_points_xy(torch.stack([w.x, w.y], dim=1).reshape(-1))
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    torch.cuda.synchronize()

    def step():
        pi, tree = cs.quadtree_on_points(pts, *bbox, w.scale, w.max_depth, w.max_size)
        bb = cs.polygon_bounding_boxes(polys)
        pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *bbox, w.scale, w.max_depth)
        res = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, polys)
        return pi, tree, pairs, res

    for i in range(args.warmup):
        log(f"[bench] warmup {i}")
        out = step()
    del out
    torch.cuda.synchronize()

    _capi.profile_reset()
    _capi.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _capi.profile_enable(False)
    prof = _capi.profile_report()
    pi, tree, pairs, res = out
    n = w.num_points
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n * args.steps / (t1 - t0)
    log(f"[bench] {ms_per_step:.3f} ms/step, {value:.4g} points/s; nodes={len(tree)} "
        f"pairs={len(pairs)} hits={len(res)}")

    # --- per-kernel algorithmic work ------------------------------------------
    key_bits = min(32, 2 * math.ceil(math.log2((1 << w.max_depth) + 3)))
    passes = math.ceil(key_bits / 8)
    ql = tree["length"].view(torch.int32).to(torch.int64)
    pq = pairs["quad_offset"].view(torch.int32).to(torch.int64)
    pp = pairs["bbox_offset"].view(torch.int32).to(torch.int64)
    edges = torch.as_tensor(effective_edges(w.part_offset, w.ring_offset, w.px, w.py), device=dev)
    cand = ql[pq]
    edge_tests = float((cand.to(torch.float64) * edges[pp]).sum())
    candidates = float(cand.sum())
    work = {
        # name: (bound, algorithmic units per launch, unit)
        # keygen reads x, y and writes the key
        "keygen": ("hbm", n * (16 + 4) / 1e9, "GB"),
        # the digit count of passes 1.. reads the keys
        "point_sort_count": ("hbm", n * 4 / 1e9, "GB"),
        # pass 0 reads keys and writes keys + indices; later passes move both
        "point_sort_scatter": ("hbm", n * (12 + 16 * (passes - 1)) / passes / 1e9, "GB"),
        # each candidate point's coordinates and index, gathered in leaf order from
        # randomly ordered input: the bound is the gather (one 128-byte line per
        # 16-byte point), not the FP64 edge arithmetic (6 flop per edge test)
        "pip_test": ("hbm", candidates * (16 + 4) / 1e9, "GB"),
    }
    kernels = {}
    for name, (cnt, tot_ms) in prof.items():
        avg_ms = tot_ms / max(cnt, 1)
        entry = {"launches_per_step": cnt / args.steps, "avg_ms": round(avg_ms, 6),
                 "ms_per_step": round(tot_ms / args.steps, 6)}
        if name in work:
            bound, units, unit = work[name]
            rate = units / (avg_ms / 1e3)
            peak = HBM_PEAK_GBS if bound == "hbm" else FP64_PEAK_TFLOPS
            entry.update(bound=bound, achieved=round(rate, 3),
                         unit="GB/s" if bound == "hbm" else "TFLOP/s",
                         frac=round(rate / peak, 4))
        if name == "pip_test":
            t = avg_ms / 1e3
            entry.update(line_granular_gbs=round(candidates * 128 / 1e9 / t, 3),
                         fp64_tflops=round(6.0 * edge_tests / 1e12 / t, 3))
        kernels[name] = entry
    dominant = max(prof.items(), key=lambda kv: kv[1][1])[0]
    dom = kernels[dominant]
    roofline = None
    if "achieved" in dom:
        bound, units, unit = work[dominant]
        roofline = {
            "kernel": dominant,
            "bound": "hbm" if bound == "hbm" else "mfma",
            "engine": "HBM" if bound == "hbm" else "FP64 VALU (peak equals the FP64 matrix peak)",
            "achieved": dom["achieved"],
            "peak": HBM_PEAK_GBS if bound == "hbm" else FP64_PEAK_TFLOPS,
            "unit": dom["unit"],
            "frac": dom["frac"],
            "traffic": measured_traffic(dominant, n),
            "avg_ms": dom["avg_ms"],
            "algorithmic_per_launch": f"{units:.6g} {unit}",
        }

    # --- CPU baseline: the oracle on a bounded prefix of the same points ---------
    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        oracle.build()
        threads = oracle.max_threads()

        def run_oracle(m):
            sub = datasets.subsample(w, m)
            t = time.perf_counter()
            tr = oracle.quadtree_on_points(sub.x, sub.y, *bbox, w.scale, w.max_depth, w.max_size)
            obb = oracle.polygon_bounding_boxes(w.part_offset, w.ring_offset, w.px, w.py)
            jb, jq = oracle.join_quadtree_and_bounding_boxes(tr[1:], obb, *bbox, w.scale,
                                                             w.max_depth)
            oracle.quadtree_point_in_polygon((jb, jq), tr[1:], tr[0], sub.x, sub.y,
                                             w.part_offset, w.ring_offset, w.px, w.py)
            return time.perf_counter() - t

        m = min(n, 1_000_000)
        dt = run_oracle(m)
        m2 = int(min(n, 60_000_000, max(m, m * args.cpu_seconds / max(dt, 1e-3))))
        if m2 > m:
            m, dt = m2, run_oracle(m2)
        cpu = {"value": m / dt, "unit": "points/s", "cores": threads, "kind": "port",
               "sample": f"first {m} points of the same workload with all "
                         f"{w.num_polygons} polygons, full path (tree, boxes, join, PIP), "
                         f"{dt:.2f} s", "seconds": round(dt, 3)}
        log(f"[bench] cpu baseline {m / dt:.4g} points/s on {threads} threads")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "gaussian_holes: plain quadtree point-in-polygon join "
                        "(tree + boxes + join + PIP), inputs resident in HBM",
            "points": n,
            "point_layout": "interleaved x, y (vec_2d, the GeoSeries storage)",
            "polygons": w.num_polygons,
            "vertices": int(w.px.size),
            "max_depth": w.max_depth,
            "max_size": w.max_size,
            "scale": w.scale,
            "seed": w.seed,
            "nodes": len(tree),
            "pairs": len(pairs),
            "candidates": candidates,
            "edge_tests": edge_tests,
            "hits": len(res),
        },
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def plain_per_job(w, polys, dev, out):
    """The plain path on the same inputs, job by job: each job's points and its own
    polygons (offsets rebased to the job, vertex arrays sliced), as a caller that
    owns one job's data passes them; built before the timed loop. Returns
    (seconds, points/s, per-job hit-count equality with the batched result)."""
    import numpy as np
    import torch

    import cuspatial_amd as cs

    n = w.num_points
    log("[bench] plain per-job calls")
    xy = w.xy
    part_h = np.asarray(w.part_offset, np.int64)
    ring_h = np.asarray(w.ring_offset, np.int64)
    ring_d = torch.as_tensor(w.ring_offset, device=dev)
    part_d = torch.as_tensor(w.part_offset, device=dev)
    jobs = []
    for j in range(w.num_jobs):
        a, b = int(w.job_point_offsets[j]), int(w.job_point_offsets[j + 1])
        q0, q1 = int(w.job_polygon_offsets[j]), int(w.job_polygon_offsets[j + 1])
        r0, r1 = int(part_h[q0]), int(part_h[q1])
        v0, v1 = int(ring_h[r0]), int(ring_h[r1])
        P = cs.Polygons((part_d[q0:q1 + 1] - r0).to(torch.int32),
                        (ring_d[r0:r1 + 1] - v0).to(torch.int32), polys.x[v0:v1], polys.y[v0:v1])
        jobs.append((cs.Points.from_points_xy(xy[2 * a:2 * b]), P, w.bboxes[j],
                     int(w.max_depths[j]), int(w.max_sizes[j]), float(w.scales[j])))
    torch.cuda.synchronize()
    t = time.perf_counter()
    plain_hits = []
    for pj, P, (x0, x1, y0, y1), d, msz, sc in jobs:
        pi, tree = cs.quadtree_on_points(pj, x0, x1, y0, y1, sc, d, msz)
        bb = cs.polygon_bounding_boxes(P)
        pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, x0, x1, y0, y1, sc, d)
        res = cs.quadtree_point_in_polygon(pairs, tree, pi, pj, P)
        plain_hits.append(len(res))
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t
    plain_value = n / plain_s
    same = [int(out.job_hit_offsets[j + 1] - out.job_hit_offsets[j]) == plain_hits[j]
            for j in range(w.num_jobs)]
    log(f"[bench] plain: {plain_s * 1e3:.1f} ms, {plain_value:.4g} points/s; hit counts equal "
        f"for {sum(same)}/{len(same)} jobs")
    return plain_s, plain_value, same


def bench_batched(args, dev):
    """Batched join: every job of the batched_jobs workload in one call per step;
    the plain path (four single-job calls per job) on the same inputs is timed
    once beside it, and every job's hits are compared between the two."""
    import numpy as np
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import _capi, datasets

    log(f"[bench] generating batched_jobs ({args.jobs} jobs) on {dev}")
    w = datasets.batched_jobs(num_jobs=args.jobs, device=dev)
    n = w.num_points
    pts = cs.Points.from_points_xy(w.xy)
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    torch.cuda.synchronize()
    log(f"[bench] {n} points, {len(polys)} polygons, {len(w.px)} vertices")

    def step():
        return cs.batched_quadtree_point_in_polygon(pts, w.job_point_offsets, polys,
                                                    w.job_polygon_offsets, w.bboxes, w.scales,
                                                    w.max_depths, w.max_sizes)

    tuned = None
    if args.autotune:
        from cuspatial_amd import autotune

        tuned = autotune.tune(step, repeats=1, log=lambda s: log("[bench] " + s))
        log(f"[bench] autotune chose {tuned['config']}: pip_test {tuned['baseline_ms']:.2f} -> "
            f"{tuned['best_ms']:.2f} ms")
    if args.graph_only:  # one eager call: the reference for the replays
        out = step()
        times, prof, ms, value = [float("nan")], {}, float("nan"), float("nan")
        plain_s = plain_value = float("nan")
        same = []
    else:
        for i in range(args.warmup):
            log(f"[bench] warmup {i}")
            out = step()
        del out
        torch.cuda.synchronize()
        _capi.profile_reset()
        _capi.profile_enable(True)
        times = []
        for i in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            out = step()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        _capi.profile_enable(False)
        prof = _capi.profile_report()
        ms = 1e3 * sum(times) / len(times)
        value = n / (ms / 1e3)
        log(f"[bench] batched: {ms:.2f} ms per step, {value:.4g} points/s, hits={out.num_hits}")
        plain_s, plain_value, same = plain_per_job(w, polys, dev, out)

    # the same call captured once as a HIP graph and replayed per step
    graph = None
    if not args.no_graph:
        t = time.perf_counter()
        plan = cs.BatchPlan(pts, w.job_point_offsets, polys, w.job_polygon_offsets, w.bboxes,
                            w.scales, w.max_depths, w.max_sizes, headroom=1.25)
        create_s = time.perf_counter() - t
        log(f"[bench] graph plan created in {create_s:.2f} s")
        for _ in range(args.warmup):
            plan.launch()
        torch.cuda.synchronize()
        gtimes = []
        for i in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            plan.launch()
            torch.cuda.synchronize()
            gtimes.append(time.perf_counter() - t)
        got = plan.result()
        identical = bool(np.array_equal(got.job_hit_offsets, out.job_hit_offsets) and all(
            torch.equal(getattr(got, k).view(torch.int32), getattr(out, k).view(torch.int32))
            for k in ("point_indices", "polygon_index", "point_index")))
        plan.close()
        gms = 1e3 * sum(gtimes) / len(gtimes)
        graph = {"ms_per_step": gms, "value": n / (gms / 1e3), "create_s": create_s,
                 "identical_to_eager": identical,
                 "p_ms": {"min": 1e3 * min(gtimes), "max": 1e3 * max(gtimes)}}
        log(f"[bench] graph: {gms:.2f} ms per step ({n / (gms / 1e3):.4g} points/s), "
            f"identical to the eager call: {identical}")
        if not identical:
            raise SystemExit("graph replay differs from the eager batched call")
    kernels = {k: {"launches_per_step": c / args.steps, "ms_per_step": round(v / args.steps, 6)}
               for k, (c, v) in prof.items()}
    eager = None if args.graph_only else {"ms_per_step": ms, "value": value}
    if graph is not None:
        value, ms = graph["value"], graph["ms_per_step"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batched_jobs: {w.num_jobs} jobs in "
                               f"{'one captured HIP graph launch' if graph else 'one batched call'}",
                   "points": n, "polygons": len(polys), "vertices": int(len(w.px)),
                   "hits": out.num_hits, "seed": w.seed},
        "plain_per_job": None if args.graph_only else {
            "value": plain_value, "unit": "points/s", "seconds": plain_s,
            "jobs_with_equal_hit_counts": sum(same)},
        "speedup_vs_plain": None if args.graph_only else value / plain_value,
        "eager_batched_call": eager,
        "graph": graph,
        "autotune": None if tuned is None else {
            "config": tuned["config"], "pip_test_ms_before": tuned["baseline_ms"],
            "pip_test_ms_after": tuned["best_ms"], "trials": len(tuned["trials"])},
        "kernels": kernels,
        "p_ms": None if args.graph_only else {"min": 1e3 * min(times), "max": 1e3 * max(times)},
    }
    print(json.dumps(line), flush=True)


def bench_learned(args, dev):
    """Learned candidate ranking (voronoi_partition: 2e8 points, 20,000 Voronoi
    cells, 256 x 256 grid): the plain quadtree join labels the first
    --train-points points, the ranker trains on them on the GPU, and then both
    the plain join and the learned locator run over all points; their pairs are
    compared exactly, and the miss rate and exact tests per point are reported
    with and without the model (list order)."""
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import _capi, datasets, ranker

    n = args.points if args.points != 100_000_000 else 200_000_000
    log(f"[bench] generating voronoi_partition with {n} points on {dev}")
    w = datasets.voronoi_partition(num_points=n, device=dev)
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    xy = torch.stack([w.x, w.y], dim=1).reshape(-1)
    del w.x, w.y
    pts = cs.Points.from_points_xy(xy)
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    t = time.perf_counter()
    grid = ranker.CandidateGrid.build(polys, 0.0, 0.0, 1e5, 256)
    grid_s = time.perf_counter() - t
    over = int((grid.counts == ranker.OVERFLOW).sum())
    log(f"[bench] {n} points, {len(polys)} polygons, grid 256x256 built in {grid_s:.2f} s "
        f"({over} cells overflow, mean candidates "
        f"{grid.counts[grid.counts <= 16].float().mean().item():.2f})")

    def plain_step():
        pi, tree = cs.quadtree_on_points(pts, *bbox, w.scale, w.max_depth, w.max_size)
        bb = cs.polygon_bounding_boxes(polys)
        pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *bbox, w.scale, w.max_depth)
        return pi, cs.quadtree_point_in_polygon(pairs, tree, pi, pts, polys)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            out = fn()
        del out
        torch.cuda.synchronize()
        times = []
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return out, 1e3 * sum(times) / len(times), times

    (pi, hits), plain_ms, _ = timed(plain_step, args.steps, args.warmup)
    plain_value = n / (plain_ms / 1e3)
    log(f"[bench] plain join: {plain_ms:.2f} ms per step, {plain_value:.4g} points/s, "
        f"{len(hits)} hits")
    pos = hits["point_index"].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    want_pt = pi.view(torch.int32).to(torch.int64)[pos] & 0xFFFFFFFF
    want_poly = hits["polygon_index"].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    del pi, hits, pos
    order = torch.argsort(want_pt * (1 << 32) + want_poly)
    want_pt, want_poly = want_pt[order], want_poly[order]
    del order
    per_point = ranker.per_point_polygon(want_poly, want_pt, n)

    # training on train_points points drawn at random, labelled by the plain join
    m = min(args.train_points, n)
    gen = torch.Generator(device=dev)
    gen.manual_seed(w.seed)
    sample = torch.randperm(n, generator=gen, device=dev)[:m]
    train_xy = xy.view(-1, 2)[sample].reshape(-1)
    labels = grid.slots_of(train_xy, per_point[sample])
    model = ranker.Ranker(grid.num_cells, seed=1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    hist = model.train(grid, train_xy, labels, epochs=args.epochs, batch=16384, lr=3e-3,
                       log=lambda s: log("[bench] train " + s))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t
    del labels, train_xy, sample

    loc = ranker.LearnedLocator(polys, grid, model, bbox, w.scale, w.max_depth, w.max_size)
    _capi.profile_reset()
    _capi.profile_enable(True)
    res, learned_ms, times = timed(lambda: loc.locate(pts), args.steps, args.warmup)
    _capi.profile_enable(False)
    prof = _capi.profile_report()
    value = n / (learned_ms / 1e3)
    poly, pt = res.pairs()
    exact = bool(torch.equal(poly, want_poly) and torch.equal(pt, want_pt))
    has = per_point >= 0
    miss = ((res.flags & 2) > 0) & has
    tests_learned = res.tests.to(torch.float64).mean().item()
    log(f"[bench] learned: {learned_ms:.2f} ms per step, {value:.4g} points/s; pairs identical "
        f"to the plain join: {exact}; miss rate {miss.sum().item() / max(has.sum().item(), 1):.4f}"
        f"; tests/point {tests_learned:.3f}; quadtree fallback {res.fallback}")
    zero = ranker.LearnedLocator(polys, grid, ranker.Ranker(grid.num_cells, zero=True), bbox,
                                 w.scale, w.max_depth, w.max_size)
    zres, zero_ms, _ = timed(lambda: zero.locate(pts), 1, 1)
    zpoly, zpt = zres.pairs()
    zexact = bool(torch.equal(zpoly, want_poly) and torch.equal(zpt, want_pt))
    zmiss = ((zres.flags & 2) > 0) & has
    tests_zero = zres.tests.to(torch.float64).mean().item()
    log(f"[bench] list order (no model): tests/point {tests_zero:.3f}, miss rate "
        f"{zmiss.sum().item() / max(has.sum().item(), 1):.4f}, pairs identical: {zexact}")
    launches = max(prof.get("ranker_rank", (1, 0.0))[0], 1)
    loc_ms = prof.get("ranker_rank", (0, 0.0))[1] / launches
    flops = 2.0 * (96 * 128 + 128 * 128 + 128 * 16) * n
    by_size = learned_vs_plain_by_batch(polys, loc, bbox, w, xy, n)
    kernels = {k: {"launches": c, "ms_per_launch": round(v / max(c, 1), 6)}
               for k, (c, v) in prof.items()}
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": learned_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 exact test; bf16 MLP with fp32 accumulation",
        "data": "synthetic",
        "config": {"workload": "voronoi_learned: point location in a Voronoi partition",
                   "points": n, "polygons": len(polys), "grid": "256x256",
                   "overflowed_cells": over, "train_points": m, "epochs": args.epochs,
                   "seed": w.seed},
        "plain_join": {"value": plain_value, "unit": "points/s", "ms_per_step": plain_ms},
        "speedup_vs_plain": value / plain_value,
        "pairs_identical_to_plain": exact,
        "miss_rate": miss.sum().item() / max(has.sum().item(), 1),
        "tests_per_point": {"learned": tests_learned, "list_order": tests_zero},
        "list_order_miss_rate": zmiss.sum().item() / max(has.sum().item(), 1),
        "quadtree_fallback_points": res.fallback,
        "training": {"seconds": train_s, "history": hist},
        "roofline": {"kernel": "rank_kernel", "bound": "mfma", "engine": "bf16 MFMA",
                     "achieved": flops / (loc_ms / 1e3) / 1e12 if loc_ms else None,
                     "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": (flops / (loc_ms / 1e3) / 1e12 / 2500.0) if loc_ms else None,
                     "traffic": None},
        "kernels": kernels,
        "p_ms": {"min": 1e3 * min(times), "max": 1e3 * max(times)},
        "by_batch_points": by_size,
    }
    print(json.dumps(line), flush=True)


def learned_vs_plain_by_batch(polys, loc, bbox, w, xy, n):
    """One request's latency by batch size, as a serving index answers it: the
    first b points of the workload located by the learned locator and by the
    plain join with the polygon boxes resident (serving.ResidentIndex.query, both
    sorted by point, then polygon). Median of 9 timed calls after 3 warm-up
    calls; the two answers are compared exactly."""
    import torch

    from cuspatial_amd import serving

    learned = serving.ResidentIndex(polys, bbox, w.scale, w.max_depth, w.max_size, locator=loc)
    plain = serving.ResidentIndex(polys, bbox, w.scale, w.max_depth, w.max_size)

    def median_ms(index, sub):
        for _ in range(3):
            index.query(sub)
        torch.cuda.synchronize()
        ts = []
        for _ in range(9):
            t0 = time.perf_counter()
            index.query(sub)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return 1e3 * sorted(ts)[len(ts) // 2]

    rows = []
    for b in (1_000, 10_000, 100_000, 1_000_000, 10_000_000, 30_000_000):
        if b > n:
            break
        sub = xy[: 2 * b]
        lms, pms = median_ms(learned, sub), median_ms(plain, sub)
        a, c = learned.query(sub), plain.query(sub)
        same = bool(all(torch.equal(u, v) for u, v in zip(a, c)))
        rows.append({"points": b, "learned_ms": lms, "plain_ms": pms, "speedup": pms / lms,
                     "identical": same})
        log(f"[bench] batch of {b} points: learned {lms:.3f} ms, plain {pms:.3f} ms "
            f"({pms / lms:.2f}x), identical: {same}")
    return rows


def bench_serving(args, dev):
    """Reverse geocoding under load: Poisson arrivals at each rate of --qps, request
    sizes log-uniform in [1, 1e6] points drawn from the Gaussian mixture; two
    resident indexes: the 50,000 star polygons of gaussian_holes (plain quadtree
    join per batch) and the 20,000-cell Voronoi partition (learned locator with
    the exact fallback; ranker trained at start-up on 1e7 points labelled by the
    plain join). Reports p50/p95/p99 latency and points/s per rate."""
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import datasets, ranker, serving

    log("[bench] serving: building the resident indexes")
    gh = datasets.gaussian_holes(num_points=1, device=dev)
    stars = cs.Polygons(gh.part_offset, gh.ring_offset, gh.px, gh.py)
    bbox = (gh.x_min, gh.x_max, gh.y_min, gh.y_max)
    plain_index = serving.ResidentIndex(stars, bbox, gh.scale, gh.max_depth, gh.max_size)
    vw = datasets.voronoi_partition(num_points=10_000_000, device=dev)
    vor = cs.Polygons(vw.part_offset, vw.ring_offset, vw.px, vw.py)
    grid = ranker.CandidateGrid.build(vor, 0.0, 0.0, 1e5, 256)
    txy = torch.stack([vw.x, vw.y], 1).reshape(-1)
    poly, pt = ranker.plain_pairs(cs.Points.from_points_xy(txy), vor, bbox, vw.scale,
                                  vw.max_depth, vw.max_size)
    labels = grid.slots_of(txy, ranker.per_point_polygon(poly, pt, txy.numel() // 2))
    model = ranker.Ranker(grid.num_cells, seed=1)
    model.train(grid, txy, labels, epochs=3, batch=16384, lr=3e-3)
    del txy, labels, poly, pt
    loc = ranker.LearnedLocator(vor, grid, model, bbox, vw.scale, vw.max_depth, vw.max_size)
    learned_index = serving.ResidentIndex(vor, bbox, vw.scale, vw.max_depth, vw.max_size,
                                          locator=loc)
    # request points: a pool of 2e8 mixture points resident on the device
    px, py = datasets._gaussian_mixture_device(2005, 200_000_000, 1e5, dev)
    pool_xy = torch.stack([px, py], 1).reshape(-1)
    del px, py
    torch.cuda.synchronize()
    rates = [float(q) for q in args.qps.split(",") if q]
    sweep = {}
    for name, index in (("plain_stars", plain_index), ("learned_voronoi", learned_index)):
        # warm the allocator and the kernels once
        serving.Server(index).serve(serving.poisson_requests(pool_xy, 1000.0, 20, seed=7))
        rows = []
        for qps in rates:
            count = max(200, int(qps * args.serve_seconds))
            reqs = serving.poisson_requests(pool_xy, qps, count, seed=int(qps))
            st = serving.Server(index, workers=3, memory_budget=32 << 30,
                                bytes_per_point=256).serve(reqs)
            row = {"qps": qps, "requests": st.requests, "points": st.points,
                   "points_per_s": st.points_per_s, "achieved_qps": st.requests / st.seconds,
                   "latency_ms": st.latency_ms, "batches": st.batches,
                   "mean_batch_requests": st.mean_batch_requests,
                   "max_in_flight": st.max_in_flight}
            rows.append(row)
            log(f"[bench] {name} qps {qps:g}: p50 {st.latency_ms['p50']:.2f} ms, p95 "
                f"{st.latency_ms['p95']:.2f} ms, p99 {st.latency_ms['p99']:.2f} ms, "
                f"{st.points_per_s:.4g} points/s, {st.batches} batches")
        sweep[name] = rows
    best = max(r["points_per_s"] for rows in sweep.values() for r in rows)
    line = {
        "metric": METRIC,
        "value": best,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson reverse-geocoding requests, sizes "
                               "log-uniform [1, 1e6] points",
                   "workers": 3, "memory_budget_gib": 32, "serve_seconds": args.serve_seconds},
        "sweep": sweep,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
