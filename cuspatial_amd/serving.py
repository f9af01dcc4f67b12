"""Reverse-geocoding request loop over a resident polygon index.

A ``ResidentIndex`` keeps the polygons (and, for a partition, the candidate grid
and trained ranker) on the device. ``Server`` runs the continuous-batching loop:

* requests (a slice of points each) arrive with their own arrival times;
* the scheduler admits waiting requests into a batch whenever a worker slot is
  free and the batch's device memory fits what the device has free: the free
  HBM that hipMemGetInfo reports, plus the bytes torch's and the library's
  caching allocators hold unused, minus the reservations of the batches in
  flight (bytes per point x points each) and a fixed reserve; an optional
  budget caps the reservations further. A request that does not fit runs alone
  once nothing else is in flight;
* each worker owns a HIP stream (default priority) and a host thread, so
  batches that are admitted while others run overlap on the device (the library
  calls release the GIL);
* a batch is answered by one query over the concatenated points: the plain
  quadtree join (quadtree_on_points -> join with the resident boxes ->
  quadtree_point_in_polygon), or the learned locator with its exact quadtree
  fallback; the pairs are split back per request and sorted by (point,
  polygon), so a request's answer does not depend on what it was batched with.

Latency = completion time - arrival time, per request.
"""

from __future__ import annotations

import collections
import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .spatial import (
    Points,
    Polygons,
    join_quadtree_and_bounding_boxes,
    polygon_bounding_boxes,
    quadtree_on_points,
    quadtree_point_in_polygon,
)


class ResidentIndex:
    """Polygons resident on the device plus the parameters each batch's quadtree
    uses (bbox, scale, max_depth, max_size). With a ``locator`` the batch is
    answered by the learned locator (partitions only).

    Without a locator and with ``captured`` (the default), a batch is answered
    by replaying a captured plan (one HIP graph launch, BatchPlan) instead of
    the four eager entry points. Plans are kept per stream and per size bucket
    (the next power of two of the batch's points, at least ``min_plan_points``):
    the batch's points are copied into the plan's point buffer and the rest of
    the bucket is filled with one point far outside the bounding box (``pad``),
    which the quadtree puts in its outside cell and which is inside no polygon,
    so the answer is the same. ``warm`` creates the plans ahead of serving. A
    batch that outgrows the
    plan's capacities (CapacityError) rebuilds the plan from its own points;
    batches above ``max_plan_points`` take the eager path (default 2^20 points:
    above it the eager path's launches and host waits are small next to the
    work, while a plan pads the batch up to its bucket's power of two, up to
    twice the points)."""

    def __init__(self, polygons: Polygons, bbox, scale, max_depth, max_size, locator=None,
                 captured: bool = True, min_plan_points: int = 1 << 12,
                 max_plan_points: int = 1 << 20, plan_headroom: float = 4.0):
        self.polygons = polygons
        self.bbox = tuple(float(v) for v in bbox)
        self.scale = float(scale)
        self.max_depth = int(max_depth)
        self.max_size = int(max_size)
        self.locator = locator
        self.boxes = polygon_bounding_boxes(polygons)  # resident across batches
        self.captured = captured and locator is None
        self.min_plan_points = int(min_plan_points)
        self.max_plan_points = int(max_plan_points)
        self.plan_headroom = float(plan_headroom)
        self._plans: dict = {}  # (stream handle, bucket) -> (BatchPlan, point buffer)
        self._plans_lock = threading.Lock()
        self.plan_stats = collections.Counter()
        # padding of a plan's point buffer: finite, outside the bbox on both axes
        x0, x1, y0, y1 = self.bbox
        self.pad = (max(x0, x1) + abs(x1 - x0) + 1.0, max(y0, y1) + abs(y1 - y0) + 1.0)

    def warm(self, xy_pool: torch.Tensor, streams, max_points: int | None = None):
        """Creates the captured plan of every size bucket up to ``max_points``
        (default ``max_plan_points``) on each of ``streams``, sized on the first
        points of ``xy_pool``, so that no request waits for a capture."""
        if not self.captured:
            return
        top = min(self.max_plan_points, max_points or self.max_plan_points,
                  xy_pool.numel() // 2)
        b = self.min_plan_points
        while b <= top:
            for s in streams:
                with torch.cuda.stream(s):
                    self._query_captured(xy_pool[:2 * b])
                s.synchronize()
            b *= 2

    def close(self):
        """Releases the captured plans."""
        with self._plans_lock:
            for plan, _ in self._plans.values():
                plan.close()
            self._plans.clear()

    def _plan_for(self, key, buf):
        from .spatial import BatchPlan

        plan = BatchPlan(Points.from_points_xy(buf), [0, buf.numel() // 2], self.polygons,
                         [0, len(self.polygons)], [self.bbox], [self.scale], [self.max_depth],
                         [self.max_size], headroom=self.plan_headroom)
        with self._plans_lock:
            self._plans[key] = (plan, buf)
        self.plan_stats["created"] += 1
        return plan

    def _query_captured(self, xy: torch.Tensor):
        from ._capi import CapacityError
        from .ranker import sort_by_point

        n = xy.numel() // 2
        bucket = max(self.min_plan_points, 1 << max(n - 1, 1).bit_length())
        if n == 0 or bucket > self.max_plan_points:
            return None
        stream = torch.cuda.current_stream(xy.device)
        key = (stream.cuda_stream, bucket)
        with self._plans_lock:
            entry = self._plans.get(key)
        if entry is None:
            buf = torch.empty(2 * bucket, dtype=torch.float64, device=xy.device)
            buf[:2 * n].copy_(xy)
            tail = buf.view(-1, 2)[n:]
            tail[:, 0].fill_(self.pad[0])
            tail[:, 1].fill_(self.pad[1])
            plan = self._plan_for(key, buf)
        else:
            plan, buf = entry
            buf[:2 * n].copy_(xy)
            tail = buf.view(-1, 2)[n:]
            tail[:, 0].fill_(self.pad[0])
            tail[:, 1].fill_(self.pad[1])
        plan.launch()
        try:
            res = plan.result()
        except CapacityError:
            plan.close()
            self.plan_stats["rebuilt"] += 1
            plan = self._plan_for(key, buf)
            plan.launch()
            res = plan.result()
        self.plan_stats["replays"] += 1
        pos = res.point_index.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        pt = res.point_indices.view(torch.int32)[pos]
        return sort_by_point(res.polygon_index.view(torch.int32), pt, n)

    def query(self, xy: torch.Tensor):
        """(polygon_index, point_index) pairs of the points in `xy`, sorted by
        point, then polygon (int64 tensors on the device)."""
        pts = Points.from_points_xy(xy)
        if self.locator is not None:
            return self.locator.locate(pts).pairs()
        if self.captured:
            out = self._query_captured(pts.xy)
            if out is not None:
                return out
        self.plan_stats["eager"] += 1
        x0, x1, y0, y1 = self.bbox
        pi, tree = quadtree_on_points(pts, x0, x1, y0, y1, self.scale, self.max_depth,
                                      self.max_size)
        pairs = join_quadtree_and_bounding_boxes(tree, self.boxes, x0, x1, y0, y1, self.scale,
                                                 self.max_depth)
        hits = quadtree_point_in_polygon(pairs, tree, pi, pts, self.polygons)
        from .ranker import join_pairs_by_point

        return join_pairs_by_point(pi, hits, len(pts))


@dataclass
class Request:
    rid: int
    xy: torch.Tensor        # interleaved points on the device
    arrival: float          # seconds from the start of the run
    done: float = math.nan  # completion time (same clock)
    result: tuple = None    # (polygon_index, point_index) local to the request

    @property
    def points(self) -> int:
        return self.xy.numel() // 2


@dataclass
class ServeStats:
    requests: int
    points: int
    seconds: float
    latency_ms: dict
    batches: int
    mean_batch_requests: float
    max_in_flight: int
    offered_qps: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def points_per_s(self) -> float:
        return self.points / self.seconds if self.seconds > 0 else 0.0


_STREAMS: dict = {}


def worker_stream(device: torch.device, k: int) -> torch.cuda.Stream:
    """Worker k's stream on `device`, created once per process (default priority):
    torch's caching allocator keeps blocks per stream, so fresh streams per run
    would start from empty caches."""
    key = (device.index if device.index is not None else torch.cuda.current_device(), k)
    if key not in _STREAMS:
        _STREAMS[key] = torch.cuda.Stream(device=device)
    return _STREAMS[key]


def free_device_bytes(device: torch.device) -> int:
    """Device memory a new batch can use now: hipMemGetInfo's free bytes plus
    what torch's caching allocator and the library's pool hold cached (both
    reuse those blocks before asking the device for more)."""
    from . import _capi

    free, _ = torch.cuda.mem_get_info(device)
    torch_cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    _, lib_cached = _capi.memory_info()
    return int(free) + int(torch_cached) + int(lib_cached)


class Server:
    """Continuous batching over a ResidentIndex with `workers` device slots.

    Admission reserves bytes_per_point per point of a batch (a bound on what the
    plain path allocates per point: sort buffers, tree, PIP workspace, outputs)
    against the device's free memory (``free_device_bytes``) less the
    reservations in flight and `reserve_bytes`; `memory_budget`, when given,
    also caps the sum of the reservations."""

    def __init__(self, index: ResidentIndex, workers: int = 3, memory_budget: int | None = None,
                 bytes_per_point: int = 256, max_batch_points: int = 1 << 24,
                 keep_results: bool = False, reserve_bytes: int = 2 << 30):
        self.index = index
        self.workers = workers
        self.memory_budget = memory_budget
        self.bytes_per_point = bytes_per_point
        self.max_batch_points = max_batch_points
        self.keep_results = keep_results
        self.reserve_bytes = reserve_bytes

    def _run_batch(self, stream, batch):
        with torch.cuda.stream(stream):
            xy = torch.cat([r.xy for r in batch]) if len(batch) > 1 else batch[0].xy
            poly, pt = self.index.query(xy)
            if self.keep_results:
                bounds = np.cumsum([0] + [r.points for r in batch])
                cut = torch.searchsorted(pt, torch.as_tensor(bounds, device=pt.device))
                cut = cut.cpu().tolist()
                for i, r in enumerate(batch):
                    a, b = cut[i], cut[i + 1]
                    r.result = (poly[a:b].clone(), pt[a:b] - int(bounds[i]))
            stream.synchronize()

    def serve(self, requests: list[Request]) -> ServeStats:
        """Runs until every request is answered; arrivals are honoured in real
        time (a request is not admitted before its arrival time)."""
        requests = sorted(requests, key=lambda r: r.arrival)
        dev = requests[0].xy.device if requests else torch.device("cuda")
        streams = [worker_stream(dev, k) for k in range(self.workers)]
        work: queue.Queue = queue.Queue()
        lock = threading.Condition()
        state = {"in_flight_bytes": 0, "in_flight": 0, "max_in_flight": 0, "done": 0,
                 "error": None, "max_in_flight_bytes": 0, "min_free_bytes": None,
                 "memory_waits": 0}
        t0 = time.perf_counter()

        def worker(k):
            while True:
                item = work.get()
                if item is None:
                    return
                batch, nbytes = item
                try:
                    self._run_batch(streams[k], batch)
                except BaseException as e:  # noqa: BLE001 - reported by serve()
                    with lock:
                        state["error"] = e
                now = time.perf_counter() - t0
                with lock:
                    for r in batch:
                        r.done = now
                    state["in_flight_bytes"] -= nbytes
                    state["in_flight"] -= 1
                    state["done"] += len(batch)
                    lock.notify_all()

        threads = [threading.Thread(target=worker, args=(k,), daemon=True)
                   for k in range(self.workers)]
        for th in threads:
            th.start()
        pending = collections.deque()
        nxt = 0
        batches = 0
        batch_sizes = []
        total = len(requests)
        try:
            with lock:
                while state["done"] < total:
                    if state["error"] is not None:
                        raise state["error"]
                    now = time.perf_counter() - t0
                    while nxt < total and requests[nxt].arrival <= now:
                        pending.append(requests[nxt])
                        nxt += 1
                    admitted = False
                    if pending and state["in_flight"] < self.workers:
                        batch, pts = [], 0
                        dev_free = free_device_bytes(dev)
                        if state["min_free_bytes"] is None or dev_free < state["min_free_bytes"]:
                            state["min_free_bytes"] = dev_free
                        free = dev_free - state["in_flight_bytes"] - self.reserve_bytes
                        if self.memory_budget is not None:
                            free = min(free, self.memory_budget - state["in_flight_bytes"])
                        while pending:
                            r = pending[0]
                            need = (pts + r.points) * self.bytes_per_point
                            if batch and (pts + r.points > self.max_batch_points or need > free):
                                break
                            if not batch and need > free and state["in_flight"] > 0:
                                state["memory_waits"] += 1
                                break  # wait for memory to free up
                            batch.append(pending.popleft())
                            pts += r.points
                        if batch:
                            nbytes = pts * self.bytes_per_point
                            state["in_flight_bytes"] += nbytes
                            state["in_flight"] += 1
                            state["max_in_flight"] = max(state["max_in_flight"],
                                                         state["in_flight"])
                            state["max_in_flight_bytes"] = max(state["max_in_flight_bytes"],
                                                               state["in_flight_bytes"])
                            work.put((batch, nbytes))
                            batches += 1
                            batch_sizes.append(len(batch))
                            admitted = True
                    if admitted:
                        continue
                    wait = None
                    if nxt < total:
                        wait = max(0.0, requests[nxt].arrival - (time.perf_counter() - t0))
                    lock.wait(timeout=wait if wait is not None else 0.05)
        finally:
            for _ in threads:
                work.put(None)
            for th in threads:
                th.join()
        elapsed = max(r.done for r in requests) - min(r.arrival for r in requests) if requests \
            else 0.0
        lat = np.array([1e3 * (r.done - r.arrival) for r in requests])
        pct = {f"p{q}": float(np.percentile(lat, q)) for q in (50, 95, 99)} if lat.size else {}
        pct["max"] = float(lat.max()) if lat.size else 0.0
        return ServeStats(total, sum(r.points for r in requests), elapsed, pct, batches,
                          float(np.mean(batch_sizes)) if batch_sizes else 0.0,
                          state["max_in_flight"],
                          extra={"max_in_flight_bytes": state["max_in_flight_bytes"],
                                 "min_free_bytes": state["min_free_bytes"],
                                 "memory_waits": state["memory_waits"]})


def poisson_requests(pool_xy: torch.Tensor, qps: float, count: int, seed: int = 0,
                     min_points: int = 1, max_points: int = 1_000_000) -> list[Request]:
    """`count` requests with exponential inter-arrival times (rate qps) and
    log-uniform sizes in [min_points, max_points], each a slice of `pool_xy`
    (interleaved points on the device)."""
    rng = np.random.default_rng(seed)
    n_pool = pool_xy.numel() // 2
    arrivals = np.cumsum(rng.exponential(1.0 / qps, count))
    sizes = np.exp(rng.uniform(math.log(min_points), math.log(max_points + 1), count))
    sizes = np.minimum(np.floor(sizes).astype(np.int64), max_points)
    sizes = np.minimum(np.maximum(sizes, min_points), n_pool)
    starts = rng.integers(0, n_pool - sizes + 1)
    return [Request(i, pool_xy[2 * s:2 * (s + k)], float(t))
            for i, (s, k, t) in enumerate(zip(starts, sizes, arrivals))]
