"""Benchmark of the quadtree point-in-polygon join on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]

The default line measures the headline metric on the batched workload (4,096
independent jobs, 5.9e8 points, 1.08M star polygons): `value` is the points/s
of the batched join replayed as one captured HIP graph per step, and
`vs_baseline` is that over the plain path on the same inputs (the four
single-job entry points, job by job, timed in the same run). Beside it, in
sub-objects of the same line:

  c1_uniform_stars    1e6 points, 1,000 polygons: the plain GPU path against the
                      CPU oracle, single-threaded and on all cores;
  c2_gaussian_holes   1e8 points, 50,000 polygons with holes: the plain path, per
                      kernel, with its roofline;
  c4_voronoi_learned  2e8 points in a 20,000-cell Voronoi partition: the learned
                      locator (ranker trained on the GPU) against the plain join
                      on the same points, miss rate, exact tests per point;
  c5_serving          reverse-geocoding requests at a fixed Poisson rate against
                      the resident 50,000 star polygons (plain join per batch)
                      and their Voronoi partition (learned locator): p50/p95/p99.

Every stage runs on the GPU through libcuspatial_amd's C ABI with inputs
resident in HBM. The roofline of the batched workload's dominant kernel comes
from the library's HIP-event timers on the eager batched call of the same batch
(a captured graph carries no per-launch events); the CPU baseline is the oracle
(the host restatement of the reference algorithm), compiled for the host CPU,
on a bounded prefix of the same jobs, single-threaded and on all cores.

--workload gaussian_holes | batched_jobs | voronoi_learned | serving |
uniform_stars prints one section alone (the section's own line).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("points/sec joined on 1 MI355X, learned+batched vs plain quadtree PIP (exact); "
          "p99 ms")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled
# by the same counters on a device copy of known size)
TRAFFIC_FILE = os.path.join(HERE, "profiles", "hbm_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sync():
    import torch

    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# accounting helpers
# ---------------------------------------------------------------------------
def effective_edges(part_offset, ring_offset, px, py):
    """Edges the crossing test visits per polygon (a torch tensor on the
    polygons' device): each ring's vertex pairs, last -> first included, minus
    zero-length ones (closed rings repeat a vertex)."""
    import torch

    ring = ring_offset.to(torch.int64)
    part = part_offset.to(torch.int64)
    dev = ring.device
    nr = ring.numel() - 1
    n = px.numel()
    prv = torch.arange(n, device=dev) - 1  # b = previous vertex, a = this vertex
    prv[ring[:-1]] = ring[1:] - 1           # the first vertex pairs with the ring's last
    zero = (px == px[prv]) & (py == py[prv])
    ring_id = torch.repeat_interleave(torch.arange(nr, device=dev), ring[1:] - ring[:-1])
    ring_edges = torch.zeros(nr, dtype=torch.int64, device=dev).index_add_(
        0, ring_id, (~zero).to(torch.int64))
    poly_of_ring = torch.repeat_interleave(torch.arange(part.numel() - 1, device=dev),
                                           part[1:] - part[:-1])
    return torch.zeros(part.numel() - 1, dtype=torch.int64, device=dev).index_add_(
        0, poly_of_ring, ring_edges)


def measured_traffic(timer: str, workload: str):
    """HBM bytes per launch of the kernel behind `timer` on `workload`, from the
    committed counter summary (rows: workload timer kernel fetch_MiB write_MiB),
    or None when it holds no such row."""
    try:
        with open(TRAFFIC_FILE) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    for ln in lines:
        fields = ln.split()
        if not fields or ln.startswith("#") or len(fields) != 5:
            continue
        if fields[0] == workload and fields[1] == timer:
            return (float(fields[3]) + float(fields[4])) * 2**20
    return None


def kernel_table(prof, steps, work):
    """Per-kernel launches and time per step from the library's timers; rate and
    fraction of peak for the kernels `work` names (name -> (bound, units per
    step, unit))."""
    kernels = {}
    for name, (cnt, tot_ms) in prof.items():
        entry = {"launches_per_step": cnt / steps, "avg_ms": round(tot_ms / max(cnt, 1), 6),
                 "ms_per_step": round(tot_ms / steps, 6)}
        if name in work:
            bound, units, unit = work[name]
            rate = units / (tot_ms / steps / 1e3)
            peak = {"hbm": HBM_PEAK_GBS, "fp64": FP64_PEAK_TFLOPS, "bf16": BF16_PEAK_TFLOPS}[bound]
            entry.update(bound=bound, achieved=round(rate, 3),
                         unit="GB/s" if bound == "hbm" else "TFLOP/s",
                         frac=round(rate / peak, 4))
        kernels[name] = entry
    return kernels


def roofline_of(kernels, name, work, workload, note):
    if name not in kernels or "achieved" not in kernels[name]:
        return None
    k = kernels[name]
    bound, units, unit = work[name]
    launches = max(k["launches_per_step"], 1e-9)
    return {
        "kernel": name,
        "bound": "hbm" if bound == "hbm" else "mfma",
        "engine": {"hbm": "HBM", "fp64": "FP64 (vector issue; the FP64 peak is the published "
                   "78.6 TFLOP/s)", "bf16": "bf16 MFMA"}[bound],
        "achieved": k["achieved"],
        "peak": {"hbm": HBM_PEAK_GBS, "fp64": FP64_PEAK_TFLOPS, "bf16": BF16_PEAK_TFLOPS}[bound],
        "unit": k["unit"],
        "frac": k["frac"],
        "traffic": measured_traffic(name, workload),
        "avg_ms": k["avg_ms"],
        "launches_per_step": k["launches_per_step"],
        "algorithmic_per_launch": f"{units / launches:.6g} {unit}",
        "measured_on": note,
    }


def host_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity": share, "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


_ORACLE = None


def oracle_native():
    """The oracle compiled -O3 -march=native for this host (once per run)."""
    global _ORACLE
    if _ORACLE is None:
        import tempfile

        import oracle

        oracle.use_native(os.path.join(tempfile.gettempdir(), f"qpip_oracle_{os.getpid()}"))
        _ORACLE = oracle
    return _ORACLE


def cpu_threads():
    o = oracle_native()
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else o.max_threads()
    try:
        n = min(n, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    return max(1, n)


def cpu_legs(run_units, seconds):
    """Times run_units(budget_s, threads) -> (units, seconds, sample description)
    single-threaded and on all cores."""
    o = oracle_native()
    out = {}
    for name, threads in (("single_thread", 1), ("all_cores", cpu_threads())):
        o.set_threads(threads)
        units, secs, sample = run_units(seconds, threads)
        out[name] = {"value": units / secs if secs > 0 else None, "cores": threads,
                     "seconds": round(secs, 3), "sample": sample}
        log(f"[bench] cpu {name} ({threads} threads): {units / max(secs, 1e-9):.4g} per s, "
            f"{sample}")
    o.set_threads(cpu_threads())
    return out


def oracle_join(o, x, y, part, ring, px, py, bbox, scale, d, ms):
    tr = o.quadtree_on_points(x, y, *bbox, scale, d, ms)
    bb = o.polygon_bounding_boxes(part, ring, px, py)
    jb, jq = o.join_quadtree_and_bounding_boxes(tr[1:], bb, *bbox, scale, d)
    return o.quadtree_point_in_polygon((jb, jq), tr[1:], tr[0], x, y, part, ring, px, py)


# ---------------------------------------------------------------------------
# C2: the plain path at 1e8 points
# ---------------------------------------------------------------------------
def plain_step_fn(cs, pts, polys, bbox, scale, d, ms):
    def step():
        pi, tree = cs.quadtree_on_points(pts, *bbox, scale, d, ms)
        bb = cs.polygon_bounding_boxes(polys)
        pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *bbox, scale, d)
        res = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, polys)
        return pi, tree, pairs, res
    return step


def section_c2(args, dev, steps, warmup, cpu_seconds):
    import numpy as np  # noqa: F401
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import _capi, datasets

    log(f"[bench] c2: gaussian_holes with {args.points} points on {dev}")
    w = datasets.gaussian_holes(num_points=args.points, device=dev)
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    pts = cs.Points.from_points_xy(torch.stack([w.x, w.y], dim=1).reshape(-1))
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    step = plain_step_fn(cs, pts, polys, bbox, w.scale, w.max_depth, w.max_size)
    for _ in range(warmup):
        out = step()
    del out
    sync()
    # the timed steps, then the same number of steps with the kernel timers on
    t0 = time.perf_counter()
    for _ in range(steps):
        out = step()
    sync()
    t1 = time.perf_counter()
    _capi.profile_reset()
    _capi.profile_enable(True)
    for _ in range(steps):
        out = step()
    sync()
    _capi.profile_enable(False)
    prof = _capi.profile_report()
    pi, tree, pairs, res = out
    n = w.num_points
    ms_per_step = (t1 - t0) * 1e3 / steps
    value = n * steps / (t1 - t0)
    # the points' PCIe transfer, for the rate of a caller whose points start in
    # (pinned) host memory: 16 bytes per point host -> device
    host_xy = torch.empty(pts.xy.numel(), dtype=torch.float64, pin_memory=True)
    host_xy.copy_(pts.xy)
    dev_xy = torch.empty_like(pts.xy)
    sync()
    t = time.perf_counter()
    for _ in range(3):
        dev_xy.copy_(host_xy, non_blocking=True)
    sync()
    h2d_ms = (time.perf_counter() - t) * 1e3 / 3
    del host_xy, dev_xy
    kernel_ms = sum(v for _, v in prof.values()) / steps
    log(f"[bench] c2: {ms_per_step:.3f} ms/step ({kernel_ms:.3f} ms in kernels), "
        f"{value:.4g} points/s; nodes={len(tree)} pairs={len(pairs)} hits={len(res)}")

    key_bits = min(32, 2 * math.ceil(math.log2((1 << w.max_depth) + 3)))
    passes = math.ceil(key_bits / 8)
    ql = tree["length"].view(torch.int32).to(torch.int64)
    pq = pairs["quad_offset"].view(torch.int32).to(torch.int64)
    pp = pairs["bbox_offset"].view(torch.int32).to(torch.int64)
    edges = effective_edges(polys.part_offset, polys.ring_offset, polys.x, polys.y)
    cand = ql[pq]
    edge_tests = float((cand.to(torch.float64) * edges[pp]).sum())
    candidates = float(cand.sum())
    work = {
        # keygen reads x, y and writes the key
        "keygen": ("hbm", n * (16 + 4) / 1e9, "GB"),
        # the digit count of passes 1.. reads the keys
        "point_sort_count": ("hbm", (passes - 1) * n * 4 / 1e9, "GB"),
        # pass 0 reads keys and writes keys + indices; later passes move both
        "point_sort_scatter": ("hbm", n * (12 + 16 * (passes - 1)) / 1e9, "GB"),
        # each candidate point's coordinates and index, gathered in leaf order from
        # randomly ordered input (one 128-byte line per 16-byte point)
        "pip_test": ("hbm", candidates * (16 + 4) / 1e9, "GB"),
    }
    kernels = kernel_table(prof, steps, work)
    t = kernels["pip_test"]["ms_per_step"] / 1e3
    kernels["pip_test"].update(line_granular_gbs=round(candidates * 128 / 1e9 / t, 3),
                               fp64_tflops=round(6.0 * edge_tests / 1e12 / t, 3))
    dominant = max(prof.items(), key=lambda kv: kv[1][1])[0]
    out_d = {
        "value": value, "unit": "points/s", "ms_per_step": ms_per_step,
        "kernel_ms_per_step": round(kernel_ms, 4),
        "with_h2d": {"value": n / ((ms_per_step + h2d_ms) / 1e3), "unit": "points/s",
                     "h2d_ms": round(h2d_ms, 3),
                     "what": "the same step plus copying the 1.6 GB of points from pinned "
                             "host memory first (not the headline: inputs are resident)"},
        "config": {"workload": "gaussian_holes: plain quadtree point-in-polygon join (tree + "
                               "boxes + join + PIP), inputs resident in HBM",
                   "points": n, "point_layout": "interleaved x, y (vec_2d)",
                   "polygons": w.num_polygons, "vertices": int(w.px.size),
                   "max_depth": w.max_depth, "max_size": w.max_size, "scale": w.scale,
                   "seed": w.seed, "nodes": len(tree), "pairs": len(pairs),
                   "candidates": candidates, "edge_tests": edge_tests, "hits": len(res)},
        "roofline": roofline_of(kernels, dominant, work, "gaussian_holes",
                                "library HIP-event timers on the plain steps"),
        "kernels": kernels,
    }
    if cpu_seconds > 0:
        o = oracle_native()

        def run(budget, threads):
            m = min(n, 100_000)
            sub = datasets.subsample(w, m)
            t = time.perf_counter()
            oracle_join(o, sub.x, sub.y, w.part_offset, w.ring_offset, w.px, w.py, bbox, w.scale,
                        w.max_depth, w.max_size)
            dt = time.perf_counter() - t
            m2 = int(min(n, 60_000_000, max(m, m * budget / max(dt, 1e-3))))
            if m2 > m:
                sub = datasets.subsample(w, m2)
                t = time.perf_counter()
                oracle_join(o, sub.x, sub.y, w.part_offset, w.ring_offset, w.px, w.py, bbox,
                            w.scale, w.max_depth, w.max_size)
                m, dt = m2, time.perf_counter() - t
            return m, dt, f"first {m} points with all {w.num_polygons} polygons, full path"

        out_d["cpu"] = cpu_legs(run, cpu_seconds)
    del w, pts, polys, out, pi, tree, pairs, res
    return out_d


# ---------------------------------------------------------------------------
# C3: the batched workload (the headline value)
# ---------------------------------------------------------------------------
def batched_sort_bytes(w):
    """Algorithmic bytes of the batched point sort's scatter passes per step:
    the library groups consecutive jobs whose composite keys (job, Morton) fit
    32 bits (the grouping restated from qpip_batch.hip's job_groups) and sorts
    each job's points on their own in a segmented radix sort of
    ceil(Morton key bits / 8) passes."""
    import numpy as np

    d = np.clip(np.asarray(w.max_depths, np.int64), 0, 15)
    cbits = np.array([int(v + 2).bit_length() for v in (1 << d)], np.int64)  # bits of 2^d + 2
    kb = np.minimum(32, 2 * cbits)
    jpo = np.asarray(w.job_point_offsets, np.int64)
    total, launches, j0, J = 0, 0, 0, len(kb)
    while j0 < J:
        j1, k = j0 + 1, int(kb[j0])
        while j1 < J:
            k2 = max(k, int(kb[j1]))
            if k2 + int(j1 - j0).bit_length() > 32 or jpo[j1 + 1] - jpo[j0] >= (1 << 31):
                break
            k, j1 = k2, j1 + 1
        n = int(jpo[j1] - jpo[j0])
        passes = max(1, math.ceil(k / 8))
        if n:
            total += n * (12 + 16 * (passes - 1))
            launches += passes
        j0 = j1
    return total, launches


def plain_per_job(w, polys, dev, out):
    """The plain path on the same inputs, job by job: each job's points and its own
    polygons (offsets rebased to the job, vertex arrays sliced), as a caller that
    owns one job's data passes them, built before the timed loop. Returns
    (seconds, points/s, per-job hit-count equality with the batched result,
    candidate points, edge tests)."""
    import numpy as np
    import torch

    import cuspatial_amd as cs

    n = w.num_points
    log("[bench] c3: plain per-job calls")
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
    sync()
    t = time.perf_counter()
    kept = []
    for pj, P, (x0, x1, y0, y1), d, msz, sc in jobs:
        pi, tree = cs.quadtree_on_points(pj, x0, x1, y0, y1, sc, d, msz)
        bb = cs.polygon_bounding_boxes(P)
        pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, x0, x1, y0, y1, sc, d)
        res = cs.quadtree_point_in_polygon(pairs, tree, pi, pj, P)
        kept.append((tree["length"], pairs, len(res)))
    sync()
    plain_s = time.perf_counter() - t
    plain_value = n / plain_s
    same = [int(out.job_hit_offsets[j + 1] - out.job_hit_offsets[j]) == kept[j][2]
            for j in range(w.num_jobs)]
    # work of the PIP test: candidate points and edge tests, from the same pairs
    edges = effective_edges(polys.part_offset, polys.ring_offset, polys.x, polys.y)
    cand = tests = 0.0
    for j, (length, pairs, _) in enumerate(kept):
        if len(pairs) == 0:
            continue
        q0 = int(w.job_polygon_offsets[j])
        c = length.view(torch.int32).to(torch.int64)[
            pairs["quad_offset"].view(torch.int32).to(torch.int64)]
        e = edges[q0 + pairs["bbox_offset"].view(torch.int32).to(torch.int64)]
        cand += float(c.sum())
        tests += float((c * e).sum())
    log(f"[bench] c3 plain: {plain_s * 1e3:.1f} ms, {plain_value:.4g} points/s; hit counts equal "
        f"for {sum(same)}/{len(same)} jobs")
    return plain_s, plain_value, same, cand, tests


def section_c3(args, dev, steps, warmup, cpu_seconds, cpu_seconds_single=None):
    import numpy as np
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import _capi, autotune, datasets

    log(f"[bench] c3: generating batched_jobs ({args.jobs} jobs) on {dev}")
    w = datasets.batched_jobs(num_jobs=args.jobs, device=dev)
    n = w.num_points
    pts = cs.Points.from_points_xy(w.xy)
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    sync()
    log(f"[bench] c3: {n} points, {len(polys)} polygons, {len(w.px)} vertices")
    edges_total = max(polys.x.numel() - (polys.ring_offset.numel() - 1), 0)
    bucket = autotune.shape_bucket(w.num_jobs, n, len(polys), edges_total)
    table = autotune.TuningTable(args.tuning_file or autotune.DEFAULT_TABLE)
    tuned = table.lookup(bucket)
    log(f"[bench] c3: shape bucket {bucket}, saved tuning {tuned}")

    def eager():
        return cs.batched_quadtree_point_in_polygon(pts, w.job_point_offsets, polys,
                                                    w.job_polygon_offsets, w.bboxes, w.scales,
                                                    w.max_depths, w.max_sizes)

    autotuned = None
    if args.autotune:
        with autotune.applied(tuned):
            autotuned = autotune.tune(eager, repeats=2, log=lambda m: log("[bench] c3 " + m))
        tuned = dict(autotuned["config"])  # `applied` restored the previous values
        if args.save_tuning:
            out_table = autotune.TuningTable(args.save_tuning if os.path.exists(args.save_tuning)
                                             else None)
            out_table.record(bucket, autotuned, "batched_jobs")
            out_table.save(args.save_tuning)
            log(f"[bench] c3: saved tuning of {bucket} to {args.save_tuning}")
    tune_ctx = autotune.applied(tuned)
    tune_ctx.__enter__()

    # the eager batched call: the reference answer and the per-kernel timers
    out = eager()
    sync()
    prof_steps = 2
    _capi.profile_reset()
    _capi.profile_enable(True)
    t = time.perf_counter()
    for _ in range(prof_steps):
        out = eager()
    sync()
    eager_ms = (time.perf_counter() - t) * 1e3 / prof_steps
    _capi.profile_enable(False)
    prof = _capi.profile_report()
    log(f"[bench] c3 eager batched call: {eager_ms:.2f} ms (timers on), hits={out.num_hits}")

    # the value: the same batch captured once as a HIP graph and replayed per step
    t = time.perf_counter()
    plan = cs.BatchPlan(pts, w.job_point_offsets, polys, w.job_polygon_offsets, w.bboxes,
                        w.scales, w.max_depths, w.max_sizes, headroom=1.25, tuning=tuned)
    create_s = time.perf_counter() - t
    for i in range(warmup):
        plan.launch()
    sync()
    t0 = time.perf_counter()
    for i in range(steps):
        plan.launch()
    sync()
    t1 = time.perf_counter()
    gms = (t1 - t0) * 1e3 / steps
    value = n * steps / (t1 - t0)
    got = plan.result()
    identical = bool(np.array_equal(got.job_hit_offsets, out.job_hit_offsets) and all(
        torch.equal(getattr(got, k).view(torch.int32), getattr(out, k).view(torch.int32))
        for k in ("point_indices", "polygon_index", "point_index")))
    plan.close()
    del got
    log(f"[bench] c3 graph: {gms:.2f} ms per step ({value:.4g} points/s), identical to the "
        f"eager call: {identical}")
    if not identical:
        raise SystemExit("graph replay differs from the eager batched call")

    tune_ctx.__exit__(None, None, None)  # the per-job baseline runs under the defaults
    plain_s, plain_value, same, candidates, edge_tests = plain_per_job(w, polys, dev, out)
    sort_bytes, sort_launches = batched_sort_bytes(w)
    work = {
        "pip_test": ("fp64", 6.0 * edge_tests / 1e12, "TFLOP"),
        "batch_sort_scatter": ("hbm", sort_bytes / 1e9, "GB"),
    }
    kernels = kernel_table(prof, prof_steps, work)
    if "pip_test" in kernels:
        tt = kernels["pip_test"]["ms_per_step"] / 1e3
        kernels["pip_test"].update(candidates=candidates, edge_tests=edge_tests,
                                   gather_gbs=round(candidates * 20 / 1e9 / tt, 3))
    dominant = max((k for k in work if k in prof), key=lambda k: prof[k][1])
    res = {
        "value": value, "ms_per_step": gms,
        "config": {"workload": f"batched_jobs: {w.num_jobs} independent jobs in one captured "
                               "HIP graph launch per step (no host synchronisation, no "
                               "allocation), inputs resident in HBM",
                   "jobs": w.num_jobs, "points": n, "polygons": len(polys),
                   "vertices": int(len(w.px)), "hits": out.num_hits, "seed": w.seed,
                   "max_depth": "ceil(log4(points / 128)) per job", "max_size": 128,
                   "shape_bucket": bucket,
                   "tuning": tuned or autotune.current(),
                   "tuning_source": ("autotuned in this run" if autotuned else
                                     "saved tuning table" if table.lookup(bucket) else
                                     "library defaults")},
        "plain_per_job": {"value": plain_value, "unit": "points/s", "seconds": plain_s,
                          "jobs_with_equal_hit_counts": sum(same), "jobs": len(same),
                          "what": "qpip_quadtree_on_points_xy -> qpip_polygon_bounding_boxes -> "
                                  "qpip_join_quadtree_and_bounding_boxes -> "
                                  "qpip_quadtree_point_in_polygon_xy per job, same inputs"},
        "graph": {"ms_per_step": gms, "create_s": round(create_s, 3),
                  "identical_to_eager": identical},
        "eager_batched_call_ms": round(eager_ms, 3),
        "roofline": roofline_of(kernels, dominant, work, "batched_jobs",
                                "library HIP-event timers on the eager batched call of the same "
                                "batch (the graph's kernels, launched one by one)"),
        "kernels": kernels,
    }
    if autotuned:
        res["autotune"] = {"pip_test_ms_before": autotuned["baseline_ms"],
                           "pip_test_ms_after": autotuned["best_ms"],
                           "trials": autotuned["trials"]}
    if cpu_seconds > 0:
        o = oracle_native()

        def run(budget, threads):
            pts_done, t_used, j = 0, 0.0, 0
            while j < w.num_jobs and t_used < budget:
                x, y = w.job_points(j)
                part, ring, px, py = w.job_polygons(j)
                b = w.bboxes[j]
                t = time.perf_counter()
                oracle_join(o, x, y, part, ring, px, py, (b[0], b[1], b[2], b[3]),
                            float(w.scales[j]), int(w.max_depths[j]), int(w.max_sizes[j]))
                t_used += time.perf_counter() - t
                pts_done += len(x)
                j += 1
            return pts_done, t_used, (f"the first {j} of {w.num_jobs} jobs ({pts_done} points), "
                                      "full path per job")

        res["cpu"] = cpu_legs(run, cpu_seconds)
    del w, pts, polys, out
    return res


# ---------------------------------------------------------------------------
# C4: learned candidate ranking over a Voronoi partition
# ---------------------------------------------------------------------------
def train_locator(cs, ranker, polys, w, xy, n, train_points, epochs, dev, log_prefix):
    """Labels train_points random points with the plain join, trains the ranker
    on the GPU and returns (locator, grid, training info, plain pairs of all n
    points or None)."""
    import torch

    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    t = time.perf_counter()
    grid = ranker.CandidateGrid.build(polys, w.x_min, w.y_min, w.x_max - w.x_min, 256)
    grid_s = time.perf_counter() - t
    m = min(train_points, n)
    gen = torch.Generator(device=dev)
    gen.manual_seed(w.seed)
    # a random training sample (with replacement: no sort on the device)
    sample = torch.randint(0, n, (m,), generator=gen, device=dev)
    train_xy = xy.view(-1, 2)[sample].reshape(-1).contiguous()
    poly, pt = ranker.plain_pairs(cs.Points.from_points_xy(train_xy), polys, bbox, w.scale,
                                  w.max_depth, w.max_size)
    labels = grid.slots_of(train_xy, ranker.per_point_polygon(poly, pt, m))
    model = ranker.Ranker(grid.num_cells, seed=1)
    sync()
    t = time.perf_counter()
    hist = model.train(grid, train_xy, labels, epochs=epochs, batch=16384, lr=3e-3,
                       log=lambda s: log(log_prefix + " train " + s))
    sync()
    train_s = time.perf_counter() - t
    t = time.perf_counter()
    loc = ranker.LearnedLocator(polys, grid, model, bbox, w.scale, w.max_depth, w.max_size)
    check_s = time.perf_counter() - t
    over = int((grid.counts == ranker.OVERFLOW).sum())
    info = {"train_points": m, "epochs": epochs, "seconds": round(train_s, 3),
            "history": hist, "grid": "256x256", "grid_build_s": round(grid_s, 3),
            "overflowed_cells": over, "partition_check_s": round(check_s, 3),
            "mean_candidates": float(grid.counts[grid.counts <= 16].float().mean().item())}
    return loc, grid, info


def section_c4(args, dev, steps, warmup, cpu_seconds, by_batch=False):
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import _capi, datasets, ranker

    n = args.learned_points
    log(f"[bench] c4: voronoi_partition with {n} points on {dev}")
    w = datasets.voronoi_partition(num_points=n, device=dev)
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    xy = torch.stack([w.x, w.y], dim=1).reshape(-1)
    pts = cs.Points.from_points_xy(xy)
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py)
    loc, grid, info = train_locator(cs, ranker, polys, w, xy, n, args.train_points, args.epochs,
                                    dev, "[bench] c4")

    def timed(fn, k, wu):
        for _ in range(wu):
            o = fn()
        del o
        sync()
        t0 = time.perf_counter()
        for _ in range(k):
            o = fn()
        sync()
        return o, (time.perf_counter() - t0) * 1e3 / k

    plain = plain_step_fn(cs, pts, polys, bbox, w.scale, w.max_depth, w.max_size)
    (pi, _, _, hits), plain_ms = timed(plain, steps, warmup)
    plain_value = n / (plain_ms / 1e3)
    want_poly, want_pt = ranker.join_pairs_by_point(pi, hits, n)  # the library's sort
    del pi, hits
    per_point = ranker.per_point_polygon(want_poly, want_pt, n)
    log(f"[bench] c4 plain join: {plain_ms:.2f} ms per step, {plain_value:.4g} points/s")

    _capi.profile_reset()
    _capi.profile_enable(True)
    res, learned_ms = timed(lambda: loc.locate(pts), steps, warmup)
    _capi.profile_enable(False)
    prof = _capi.profile_report()
    value = n / (learned_ms / 1e3)
    poly, pt = res.pairs()
    exact = bool(torch.equal(poly, want_poly) and torch.equal(pt, want_pt))
    has = per_point >= 0
    miss = ((res.flags & 2) > 0) & has
    tests_learned = res.tests.to(torch.float64).mean().item()
    zero = ranker.LearnedLocator(polys, grid, ranker.Ranker(grid.num_cells, zero=True), bbox,
                                 w.scale, w.max_depth, w.max_size, check_partition=False)
    zres, _ = timed(lambda: zero.locate(pts), 1, 1)
    zpoly, zpt = zres.pairs()
    zexact = bool(torch.equal(zpoly, want_poly) and torch.equal(zpt, want_pt))
    zmiss = ((zres.flags & 2) > 0) & has
    tests_zero = zres.tests.to(torch.float64).mean().item()
    launches = max(prof.get("ranker_rank", (1, 0.0))[0], 1)
    work = {"ranker_rank": ("bf16", 2.0 * (96 * 128 + 128 * 128 + 128 * 16) * n / 1e12, "TFLOP")}
    kernels = kernel_table(prof, launches, work)
    log(f"[bench] c4 learned: {learned_ms:.2f} ms per step, {value:.4g} points/s; identical: "
        f"{exact}; miss rate {miss.sum().item() / max(has.sum().item(), 1):.4f}; tests/point "
        f"{tests_learned:.3f} (list order {tests_zero:.3f})")
    out = {
        "value": value, "unit": "points/s", "ms_per_step": learned_ms,
        "plain_join": {"value": plain_value, "ms_per_step": plain_ms},
        "speedup_vs_plain": value / plain_value,
        "pairs_identical_to_plain": exact,
        "list_order_pairs_identical_to_plain": zexact,
        "miss_rate": miss.sum().item() / max(has.sum().item(), 1),
        "list_order_miss_rate": zmiss.sum().item() / max(has.sum().item(), 1),
        "tests_per_point": {"learned": tests_learned, "list_order": tests_zero},
        "quadtree_fallback_points": res.fallback,
        "config": {"workload": "voronoi_learned: point location in a clipped Voronoi "
                               "partition (learned locator, exact FP64 confirmation, quadtree "
                               "fallback) against the plain join on the same points",
                   "points": n, "polygons": len(polys), "seed": w.seed},
        "training": info,
        "roofline": roofline_of(kernels, "ranker_rank", work, "voronoi_learned",
                                "library HIP-event timers on the timed locate calls"),
        "kernels": kernels,
    }
    if by_batch:
        out["by_batch_points"] = learned_vs_plain_by_batch(polys, loc, bbox, w, xy, n)
    if cpu_seconds > 0:
        o = oracle_native()
        hx, hy = w.x, w.y

        def run(budget, threads):
            m = 100_000
            t = time.perf_counter()
            oracle_join(o, hx[:m].cpu().numpy(), hy[:m].cpu().numpy(), w.part_offset,
                        w.ring_offset, w.px, w.py, bbox, w.scale, w.max_depth, w.max_size)
            dt = time.perf_counter() - t
            m2 = int(min(n, 60_000_000, max(m, m * budget / max(dt, 1e-3))))
            if m2 > m:
                x, y = hx[:m2].cpu().numpy(), hy[:m2].cpu().numpy()
                t = time.perf_counter()
                oracle_join(o, x, y, w.part_offset, w.ring_offset, w.px, w.py, bbox, w.scale,
                            w.max_depth, w.max_size)
                m, dt = m2, time.perf_counter() - t
            return m, dt, f"first {m} points with all {len(polys)} cells, plain full path"

        out["cpu"] = cpu_legs(run, cpu_seconds)
    return out


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
        sync()
        ts = []
        for _ in range(9):
            t0 = time.perf_counter()
            index.query(sub)
            sync()
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


# ---------------------------------------------------------------------------
# C5: serving
# ---------------------------------------------------------------------------
def section_c5(args, dev, rates):
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import datasets, ranker, serving

    log("[bench] c5: building the resident indexes")
    gh = datasets.gaussian_holes(num_points=1, device=dev)
    stars = cs.Polygons(gh.part_offset, gh.ring_offset, gh.px, gh.py)
    bbox = (gh.x_min, gh.x_max, gh.y_min, gh.y_max)
    plain_index = serving.ResidentIndex(stars, bbox, gh.scale, gh.max_depth, gh.max_size)
    t = time.perf_counter()
    vw = datasets.gaussian_holes_partition(num_points=args.train_points, device=dev)
    vor_s = time.perf_counter() - t
    vor = cs.Polygons(vw.part_offset, vw.ring_offset, vw.px, vw.py)
    txy = torch.stack([vw.x, vw.y], 1).reshape(-1)
    loc, grid, info = train_locator(cs, ranker, vor, vw, txy, txy.numel() // 2,
                                    args.train_points, args.epochs, dev, "[bench] c5")
    info["voronoi_s"] = round(vor_s, 3)
    del txy
    learned_index = serving.ResidentIndex(vor, bbox, vw.scale, vw.max_depth, vw.max_size,
                                          locator=loc)
    # request points: a pool of Gaussian-mixture points resident on the device
    px, py = datasets._gaussian_mixture_device(2005, args.serve_pool, 1e5, dev)
    pool_xy = torch.stack([px, py], 1).reshape(-1)
    del px, py
    sync()
    sweep = {}
    for name, index in (("plain_stars", plain_index), ("learned_voronoi", learned_index)):
        # warm the allocators and the kernels once, and capture the plain index's
        # plans of every batch size bucket on every worker stream
        serving.Server(index).serve(serving.poisson_requests(pool_xy, 1000.0, 20, seed=7))
        t = time.perf_counter()
        index.warm(pool_xy, [serving.worker_stream(dev, k) for k in range(3)])
        warm_s = time.perf_counter() - t
        index.plan_stats.clear()
        rows = []
        for qps in rates:
            count = max(200, int(qps * args.serve_seconds))
            reqs = serving.poisson_requests(pool_xy, qps, count, seed=int(qps))
            st = serving.Server(index, workers=3).serve(reqs)
            row = {"qps": qps, "requests": st.requests, "points": st.points,
                   "points_per_s": st.points_per_s, "achieved_qps": st.requests / st.seconds,
                   "latency_ms": st.latency_ms, "batches": st.batches,
                   "mean_batch_requests": st.mean_batch_requests,
                   "max_in_flight": st.max_in_flight, "admission": st.extra,
                   "plans": dict(index.plan_stats), "plan_warm_s": round(warm_s, 3)}
            index.plan_stats.clear()
            rows.append(row)
            log(f"[bench] c5 {name} qps {qps:g}: p50 {st.latency_ms['p50']:.2f} ms, p95 "
                f"{st.latency_ms['p95']:.2f} ms, p99 {st.latency_ms['p99']:.2f} ms, "
                f"{st.points_per_s:.4g} points/s, {st.batches} batches")
        sweep[name] = rows
    return {"config": {"workload": "serving: Poisson reverse-geocoding requests, sizes "
                                   "log-uniform [1, 1e6] points from the Gaussian mixture",
                       "indexes": {"plain_stars": "gaussian_holes' 50,000 star polygons, "
                                                  "plain quadtree join per batch",
                                   "learned_voronoi": "the clipped Voronoi cells of the same "
                                                      "50,000 polygon centres, learned locator "
                                                      "with exact fallback"},
                       "workers": 3, "serve_seconds": args.serve_seconds,
                       "admission": "free HBM (hipMemGetInfo plus the allocators' cached "
                                    "bytes) minus the reservations of batches in flight"},
            "training": info, "sweep": sweep}


# ---------------------------------------------------------------------------
# C1: the CPU-runnable configuration
# ---------------------------------------------------------------------------
def section_c1(args, dev, steps, warmup, cpu_seconds):
    import numpy as np
    import torch

    import cuspatial_amd as cs
    from cuspatial_amd import datasets

    w = datasets.uniform_stars()
    bbox = (w.x_min, w.x_max, w.y_min, w.y_max)
    xy = torch.as_tensor(np.stack([w.x, w.y], 1).reshape(-1), device=dev)
    pts = cs.Points.from_points_xy(xy)
    polys = cs.Polygons(w.part_offset, w.ring_offset, w.px, w.py, device=dev)
    step = plain_step_fn(cs, pts, polys, bbox, w.scale, w.max_depth, w.max_size)
    for _ in range(warmup):
        out = step()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = step()
    sync()
    gms = (time.perf_counter() - t0) * 1e3 / steps
    pi, tree, pairs, res = out
    res_d = {"gpu": {"value": w.num_points / (gms / 1e3), "unit": "points/s",
                     "ms_per_step": gms, "hits": len(res)},
             "config": w.describe()}
    if cpu_seconds > 0:
        o = oracle_native()
        holder = {}

        def run(budget, threads):
            t = time.perf_counter()
            holder["r"] = oracle_join(o, w.x, w.y, w.part_offset, w.ring_offset, w.px, w.py,
                                      bbox, w.scale, w.max_depth, w.max_size)
            return w.num_points, time.perf_counter() - t, f"all {w.num_points} points"

        res_d["cpu"] = cpu_legs(run, cpu_seconds)
        opoly, opt = holder["r"]
        res_d["gpu_equals_cpu"] = bool(
            np.array_equal(res["polygon_index"].cpu().numpy().astype(np.int64),
                           np.asarray(opoly, np.int64)) and
            np.array_equal(res["point_index"].cpu().numpy().astype(np.int64),
                           np.asarray(opt, np.int64)))
        for k, v in res_d["cpu"].items():
            if v["value"]:
                res_d[f"gpu_vs_cpu_{k}"] = res_d["gpu"]["value"] / v["value"]
    return res_d


# ---------------------------------------------------------------------------
def base_line(value, unit, steps, warmup, ms, config, dtype="f64"):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def north_star(args, dev):
    sections = set(args.sections.split(","))
    c3 = section_c3(args, dev, args.steps, args.warmup, args.cpu_seconds)
    line = base_line(c3["value"], "points/s", args.steps, args.warmup, c3["ms_per_step"],
                     c3["config"])
    line["vs_baseline"] = c3["value"] / c3["plain_per_job"]["value"]
    line["baseline"] = ("the plain quadtree path (four single-job entry points per job) on the "
                        "same 4,096 jobs, timed in this run: vs_baseline = value / "
                        "plain_per_job.value")
    line["roofline"] = c3["roofline"]
    cpu = c3.get("cpu")
    if cpu:
        line["cpu_baseline"] = dict(cpu["all_cores"], unit="points/s", kind="port",
                                    single_thread=cpu["single_thread"], **host_info())
    else:
        line["cpu_baseline"] = None
    line["learned_batched"] = (
        "learned ranking applies to non-overlapping polygons (c4, c5); this workload's star "
        "polygons overlap, so its value is the batched path alone")
    for k in ("plain_per_job", "graph", "eager_batched_call_ms", "kernels"):
        line[k] = c3[k]
    other = args.cpu_seconds_other if args.cpu_seconds > 0 else 0.0
    if "c2" in sections:
        line["c2_gaussian_holes"] = section_c2(args, dev, args.steps, args.warmup, other)
        sync()
    if "c4" in sections:
        line["c4_voronoi_learned"] = section_c4(args, dev, max(3, args.steps // 4), 1, other)
        sync()
    if "c5" in sections:
        c5 = section_c5(args, dev, [float(q) for q in args.qps.split(",") if q])
        line["c5_serving"] = c5
        line["p99_ms"] = {name: {"qps": rows[0]["qps"], "p99": rows[0]["latency_ms"]["p99"]}
                          for name, rows in c5["sweep"].items()}
    if "c1" in sections:
        line["c1_uniform_stars"] = section_c1(args, dev, args.steps, args.warmup,
                                              other if args.cpu_seconds > 0 else 0.0)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="north_star",
                    choices=["north_star", "gaussian_holes", "batched_jobs", "voronoi_learned",
                             "serving", "uniform_stars"])
    ap.add_argument("--sections", default="c2,c4,c5,c1",
                    help="north_star: sections printed beside the batched value")
    ap.add_argument("--points", type=int, default=100_000_000, help="gaussian_holes points")
    ap.add_argument("--learned-points", type=int, default=200_000_000,
                    help="voronoi_learned points")
    ap.add_argument("--jobs", type=int, default=4096, help="jobs of the batched workload")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU work per leg of the baseline on the value's workload (0: none)")
    ap.add_argument("--cpu-seconds-other", type=float, default=4.0,
                    help="CPU work per leg of the other sections' baselines")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--qps", default="10000", help="request rates of the serving section")
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="arrival window per rate of the serving section")
    ap.add_argument("--serve-pool", type=int, default=100_000_000,
                    help="points in the request pool of the serving section")
    ap.add_argument("--train-points", type=int, default=10_000_000,
                    help="labelled points the ranker is trained on")
    ap.add_argument("--epochs", type=int, default=3, help="ranker training epochs")
    ap.add_argument("--by-batch", action="store_true",
                    help="voronoi_learned: also time single requests by batch size")
    ap.add_argument("--tuning-file", default=None,
                    help="batched workload: the tuning table to look the batch's shape up in "
                         "(default: the package's cuspatial_amd/tuning.json)")
    ap.add_argument("--autotune", action="store_true",
                    help="batched workload: run the autotuner on the batch before timing")
    ap.add_argument("--save-tuning", default=None, metavar="PATH",
                    help="with --autotune: record the result in the tuning table at PATH")
    ap.add_argument("--tune", default="", metavar="NAME=VALUE,...",
                    help="set library tuning parameters for the whole run (a saved tuning "
                         "table still applies its own values to the batched workload)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (the known "
                         "traffic that scales the HBM counters) and exit")
    args = ap.parse_args()
    if args.no_cpu_baseline:
        args.cpu_seconds = 0.0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the path has no "
              "multi-GPU mode in this build", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.tune:
        from cuspatial_amd import autotune

        for item in args.tune.split(","):
            name, value = item.split("=")
            autotune.set(name.strip(), int(value))
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        log(f"[bench] copied {args.copy_calibration} MiB three times")
        return
    wl = args.workload
    if wl == "north_star":
        line = north_star(args, dev)
    elif wl == "batched_jobs":
        c3 = section_c3(args, dev, args.steps, args.warmup, args.cpu_seconds)
        line = base_line(c3["value"], "points/s", args.steps, args.warmup, c3["ms_per_step"],
                         c3["config"])
        line["vs_baseline"] = c3["value"] / c3["plain_per_job"]["value"]
        line.update({k: v for k, v in c3.items() if k not in ("value", "ms_per_step", "config")})
    elif wl == "gaussian_holes":
        c2 = section_c2(args, dev, args.steps, args.warmup, args.cpu_seconds)
        line = base_line(c2["value"], "points/s", args.steps, args.warmup, c2["ms_per_step"],
                         c2["config"])
        line.update({k: v for k, v in c2.items() if k not in ("value", "ms_per_step", "config")})
    elif wl == "voronoi_learned":
        c4 = section_c4(args, dev, args.steps, args.warmup, args.cpu_seconds, args.by_batch)
        line = base_line(c4["value"], "points/s", args.steps, args.warmup, c4["ms_per_step"],
                         c4["config"], dtype="f64 exact test; bf16 MLP with fp32 accumulation")
        line.update({k: v for k, v in c4.items() if k not in ("value", "ms_per_step", "config")})
    elif wl == "serving":
        c5 = section_c5(args, dev, [float(q) for q in args.qps.split(",") if q])
        best = max(r["points_per_s"] for rows in c5["sweep"].values() for r in rows)
        line = base_line(best, "points/s", 1, 1, None, c5["config"])
        line.update({k: v for k, v in c5.items() if k != "config"})
    else:
        c1 = section_c1(args, dev, args.steps, args.warmup, args.cpu_seconds)
        line = base_line(c1["gpu"]["value"], "points/s", args.steps, args.warmup,
                         c1["gpu"]["ms_per_step"], c1["config"])
        line.update({k: v for k, v in c1.items() if k != "config"})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
