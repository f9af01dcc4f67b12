"""Reverse-geocoding request loop over a resident polygon index.

A ``ResidentIndex`` keeps the polygons (and, for a partition, the candidate grid
and trained ranker) on the device. ``Server`` runs the continuous-batching loop:

* requests (a slice of points each) arrive with their own arrival times;
* the scheduler admits waiting requests into a batch whenever a worker slot is
  free and the batch's device memory fits the budget (bytes per point in
  flight x points, summed over the batches on the device); a request larger
  than the whole budget runs alone once nothing else is in flight;
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
    answered by the learned locator (partitions only)."""

    def __init__(self, polygons: Polygons, bbox, scale, max_depth, max_size, locator=None):
        self.polygons = polygons
        self.bbox = tuple(float(v) for v in bbox)
        self.scale = float(scale)
        self.max_depth = int(max_depth)
        self.max_size = int(max_size)
        self.locator = locator
        self.boxes = polygon_bounding_boxes(polygons)  # resident across batches

    def query(self, xy: torch.Tensor):
        """(polygon_index, point_index) pairs of the points in `xy`, sorted by
        point, then polygon (int64 tensors on the device)."""
        pts = Points.from_points_xy(xy)
        if self.locator is not None:
            return self.locator.locate(pts).pairs()
        x0, x1, y0, y1 = self.bbox
        pi, tree = quadtree_on_points(pts, x0, x1, y0, y1, self.scale, self.max_depth,
                                      self.max_size)
        pairs = join_quadtree_and_bounding_boxes(tree, self.boxes, x0, x1, y0, y1, self.scale,
                                                 self.max_depth)
        hits = quadtree_point_in_polygon(pairs, tree, pi, pts, self.polygons)
        pos = hits["point_index"].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        pt = pi.view(torch.int32).to(torch.int64)[pos] & 0xFFFFFFFF
        poly = hits["polygon_index"].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        order = torch.argsort(pt * (1 << 32) + poly)
        return poly[order], pt[order]


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


class Server:
    """Continuous batching over a ResidentIndex with `workers` device slots."""

    def __init__(self, index: ResidentIndex, workers: int = 3, memory_budget: int = 8 << 30,
                 bytes_per_point: int = 256, max_batch_points: int = 1 << 24,
                 keep_results: bool = False):
        self.index = index
        self.workers = workers
        self.memory_budget = memory_budget
        self.bytes_per_point = bytes_per_point
        self.max_batch_points = max_batch_points
        self.keep_results = keep_results

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
                 "error": None, "max_in_flight_bytes": 0}
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
                        free = self.memory_budget - state["in_flight_bytes"]
                        while pending:
                            r = pending[0]
                            need = (pts + r.points) * self.bytes_per_point
                            if batch and (pts + r.points > self.max_batch_points or need > free):
                                break
                            if not batch and need > free and state["in_flight"] > 0:
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
                          extra={"max_in_flight_bytes": state["max_in_flight_bytes"]})


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
