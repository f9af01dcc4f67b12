"""Autotuner for the library's launch parameters.

The tunables are the library's runtime tuning parameters (qpip_tuning_set /
qpip_tuning_get in include/cuspatial_amd.h); results never depend on them, only
the time. ``tune`` runs a caller-supplied step (one pass of the real workload)
under each candidate value, one parameter at a time (coordinate descent from the
current values), timing the kernels the parameter affects with the library's
HIP-event timers, and keeps the fastest. The chosen values are applied, and can
be saved to and loaded from a JSON file next to the workload that chose them.

A ``TuningTable`` keeps the best values per shape bucket (number of jobs,
points per job, edges per polygon, each rounded to a power of two) in a JSON
file; ``BatchPlan`` looks its shape up in the package's table (``tuning.json``,
written by ``python bench.py --workload batched_jobs --autotune --save-tuning``)
and captures its launch sequence under those values.
"""

from __future__ import annotations

import contextlib
import ctypes
import json
import math
import os
import time

from . import _capi

# parameter -> candidates tried by default, and the timer(s) it is judged by
PARAMETERS = {
    "pip_min_pairs_per_wave": ([1, 2, 4, 8, 16, 32, 64, 256], ("pip_test",)),
    "pip_pairs_shift": ([10, 12, 14, 16, 18, 20, 24], ("pip_test",)),
    "pip_waves_per_block": ([2, 4, 8], ("pip_test",)),
    "pip_stage_edges": ([64, 128], ("pip_test",)),
    "pip_xcd_run": ([1, 4, 16, 64], ("pip_test",)),
    "pip_box_filter": ([0, 1], ("pip_test",)),
    "pip_polygon_group": ([0, 4, 16, 64], ("pip_test", "pip_order")),
}


def get(name: str) -> int:
    v = ctypes.c_int64()
    _capi.check(_capi.lib().qpip_tuning_get(name.encode(), ctypes.byref(v)))
    return int(v.value)


def set(name: str, value: int) -> None:  # noqa: A001 - mirrors qpip_tuning_set
    _capi.check(_capi.lib().qpip_tuning_set(name.encode(), int(value)))


def current() -> dict:
    return {k: get(k) for k in PARAMETERS}


def apply(config: dict) -> None:
    for k, v in config.items():
        set(k, v)


def _measure(step, timers, repeats):
    import torch

    step()  # warm: allocations, caches
    torch.cuda.synchronize()
    _capi.profile_reset()
    _capi.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(repeats):
        step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / repeats
    _capi.profile_enable(False)
    prof = _capi.profile_report()
    kern = sum(prof.get(t, (0, 0.0))[1] for t in timers) / repeats
    return kern, 1e3 * wall


def tune(step, parameters=None, repeats: int = 2, rounds: int = 1, log=None) -> dict:
    """Coordinate descent over `parameters` (name -> candidates; default
    PARAMETERS). Returns {"config": best values, "trials": [...], "baseline_ms",
    "best_ms"}; the best values are left applied."""
    params = parameters or {k: v[0] for k, v in PARAMETERS.items()}
    timers = sorted({t for k in params for t in PARAMETERS.get(k, ([], ("pip_test",)))[1]})
    best = current()
    base_kern, base_wall = _measure(step, timers, repeats)
    best_kern = base_kern
    trials = [{"config": dict(best), "kernel_ms": base_kern, "step_ms": base_wall}]
    if log:
        log(f"autotune: start {best} kernel {base_kern:.3f} ms, step {base_wall:.3f} ms")
    for _ in range(rounds):
        for name, cands in params.items():
            for v in cands:
                if v == best[name]:
                    continue
                cfg = dict(best, **{name: v})
                apply(cfg)
                kern, wall = _measure(step, timers, repeats)
                trials.append({"config": cfg, "kernel_ms": kern, "step_ms": wall})
                if log:
                    log(f"autotune: {name}={v}: kernel {kern:.3f} ms, step {wall:.3f} ms")
                if kern < best_kern:
                    best, best_kern = cfg, kern
            apply(best)
    return {"config": best, "timers": timers, "baseline_ms": base_kern, "best_ms": best_kern,
            "trials": trials}


def save(path, result: dict, workload: str) -> None:
    with open(path, "w") as f:
        json.dump({"workload": workload, "config": result["config"],
                   "baseline_ms": result["baseline_ms"], "best_ms": result["best_ms"]}, f,
                  indent=1)


def load(path, bucket: str | None = None) -> dict:
    """Applies a saved configuration and returns it: the file of ``save``, or
    the entry of `bucket` in a TuningTable file (nothing applied when the table
    has no such bucket)."""
    with open(path) as f:
        data = json.load(f)
    if "buckets" in data:
        cfg = TuningTable(path).lookup(bucket) if bucket else None
        if cfg:
            apply(cfg)
        return cfg or {}
    cfg = data["config"]
    apply(cfg)
    return cfg


DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def shape_bucket(num_jobs: int, num_points: int, num_polygons: int, num_edges: int) -> str:
    """The bucket of a batch's shape: log2 of the job count, of the mean points
    per job and of the mean edges per polygon, each rounded."""
    def lg(v):
        return int(round(math.log2(max(float(v), 1.0))))

    return (f"jobs{lg(num_jobs)}_points{lg(num_points / max(num_jobs, 1))}"
            f"_edges{lg(num_edges / max(num_polygons, 1))}")


@contextlib.contextmanager
def applied(config: dict | None):
    """Runs the block under `config` and restores the previous values after."""
    if not config:
        yield
        return
    prev = current()
    apply(config)
    try:
        yield
    finally:
        apply(prev)


class TuningTable:
    """Best tuning values per shape bucket, kept in a JSON file:
    {"version": 1, "buckets": {bucket: {"config": {...}, "workload": ...,
    "baseline_ms": ..., "best_ms": ...}}}."""

    VERSION = 1

    def __init__(self, path: str | None = None):
        self.path = path
        self.buckets: dict = {}
        if path and os.path.exists(path):
            with open(path) as f:
                data = json.load(f)
            if data.get("version") != self.VERSION:
                raise ValueError(f"{path}: tuning table version {data.get('version')} is not "
                                 f"{self.VERSION}")
            self.buckets = data.get("buckets", {})
            for b in self.buckets.values():
                unknown = {*b["config"]} - {*PARAMETERS}  # `set` is qpip_tuning_set here
                if unknown:
                    raise ValueError(f"{path}: unknown tuning parameters {sorted(unknown)}")

    def lookup(self, bucket: str) -> dict | None:
        entry = self.buckets.get(bucket)
        return dict(entry["config"]) if entry else None

    def record(self, bucket: str, result: dict, workload: str) -> None:
        self.buckets[bucket] = {"config": dict(result["config"]), "workload": workload,
                                "baseline_ms": result["baseline_ms"],
                                "best_ms": result["best_ms"]}

    def save(self, path: str | None = None) -> str:
        path = path or self.path
        with open(path + ".tmp", "w") as f:
            json.dump({"version": self.VERSION, "buckets": self.buckets}, f, indent=1,
                      sort_keys=True)
        os.replace(path + ".tmp", path)
        return path
