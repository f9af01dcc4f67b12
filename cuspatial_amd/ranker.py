"""Learned candidate ranking for point location over non-overlapping polygons.

A ``CandidateGrid`` lists, for each cell of a dim x dim grid over the polygons'
domain, every polygon whose bounding box reaches the cell (up to 16; a cell with
more sends its points to the quadtree path). A ``Ranker`` (a 96-128-128-16 MLP
on the cell's learned embedding and Fourier features of the point's position in
the cell, bf16 on MFMA with fp32 accumulation) orders those candidates, and
``LearnedLocator`` tests them in that order with the quadtree path's exact FP64
point-in-polygon test, stopping at the first confirmed hit. Points that the
locator cannot finish exactly go through the plain quadtree path
(quadtree_on_points -> join -> quadtree_point_in_polygon), which is also what
labels the training points and judges the model's miss rate.

The model is trained on the GPU by the library's own forward and backward
kernels (``qpip_ranker_gradients``) and Adam (``qpip_ranker_adam``).

For partitions (polygons that do not overlap) every point is in at most one
polygon, and the locator's per-point answer, as (polygon_index, point_index)
pairs, equals the plain join's pairs exactly.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from .spatial import (
    Points,
    Polygons,
    join_quadtree_and_bounding_boxes,
    polygon_bounding_boxes,
    quadtree_on_points,
    quadtree_point_in_polygon,
)

MAX_CANDIDATES = 16
OVERFLOW = 255  # count of a cell with more than MAX_CANDIDATES candidates
EMB, FOUR, IN, HID = 64, 32, 96, 128


class qpip_grid(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("cell", ctypes.c_double),
                ("dim", ctypes.c_int32), ("candidates", ctypes.c_void_p),
                ("counts", ctypes.c_void_p)]


class qpip_ranker(ctypes.Structure):
    _fields_ = [("params", ctypes.c_void_p), ("params_bf16", ctypes.c_void_p),
                ("num_cells", ctypes.c_int64)]


def _grid_coord_np(v, v0, cell, dim):
    t = np.floor((np.asarray(v, np.float64) - v0) * (1.0 / cell))
    t = np.where(t >= 0.0, t, 0.0)
    return np.minimum(t, dim - 1).astype(np.int64)


def _grid_coord_t(v: torch.Tensor, v0, cell, dim) -> torch.Tensor:
    t = torch.floor((v - v0) * (1.0 / cell))
    t = torch.where(t >= 0.0, t, torch.zeros_like(t))
    return torch.clamp(t, max=dim - 1).to(torch.int64)


@dataclass
class CandidateGrid:
    x0: float
    y0: float
    cell: float
    dim: int
    candidates: torch.Tensor  # int32 [dim * dim * 16], -1 padded, ascending polygon id
    counts: torch.Tensor      # uint8 [dim * dim]; OVERFLOW when more than 16

    @classmethod
    def build(cls, polygons: Polygons, x0: float, y0: float, extent: float, dim: int = 256,
              device=None) -> "CandidateGrid":
        """Lists every polygon whose bounding box reaches each cell. The cell of a
        coordinate is floor((v - v0) * (1 / cell)) clamped to [0, dim - 1], the
        same function the kernels use, so the list of a point's cell contains every
        polygon that contains the point."""
        dev = device if device is not None else polygons.device
        part = polygons.part_offset.cpu().numpy().astype(np.int64)
        ring = polygons.ring_offset.cpu().numpy().astype(np.int64)
        px = polygons.x.cpu().numpy()
        py = polygons.y.cpu().numpy()
        cell = extent / dim
        vstart = ring[part[:-1]]
        vend = ring[part[1:]]
        nonempty = np.nonzero(vend > vstart)[0]
        idx = vstart[nonempty]
        mnx = np.minimum.reduceat(px, idx) if idx.size else np.zeros(0)
        mxx = np.maximum.reduceat(px, idx) if idx.size else np.zeros(0)
        mny = np.minimum.reduceat(py, idx) if idx.size else np.zeros(0)
        mxy = np.maximum.reduceat(py, idx) if idx.size else np.zeros(0)
        # reduceat runs to the next start; the last run to the array's end: trim
        # runs that extend past each polygon's own vertices
        for arr, f, src in ((mnx, np.min, px), (mxx, np.max, px), (mny, np.min, py),
                            (mxy, np.max, py)):
            nxt = np.append(idx[1:], len(src))
            bad = np.nonzero(nxt != vend[nonempty])[0]
            for b in bad:
                arr[b] = f(src[vstart[nonempty[b]]:vend[nonempty[b]]])
        cx0 = _grid_coord_np(mnx, x0, cell, dim)
        cx1 = _grid_coord_np(mxx, x0, cell, dim)
        cy0 = _grid_coord_np(mny, y0, cell, dim)
        cy1 = _grid_coord_np(mxy, y0, cell, dim)
        w = cx1 - cx0 + 1
        h = cy1 - cy0 + 1
        k = w * h
        poly = np.repeat(nonempty, k)
        j = np.arange(k.sum()) - np.repeat(np.cumsum(k) - k, k)
        ww = np.repeat(w, k)
        cells = (np.repeat(cy0, k) + j // ww) * dim + np.repeat(cx0, k) + j % ww
        order = np.lexsort((poly, cells))
        cells, poly = cells[order], poly[order]
        ncell = dim * dim
        cnt = np.bincount(cells, minlength=ncell)
        first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        slot = np.arange(cells.size) - first[cells]
        keep = slot < MAX_CANDIDATES
        cand = np.full(ncell * MAX_CANDIDATES, -1, np.int32)
        cand[cells[keep] * MAX_CANDIDATES + slot[keep]] = poly[keep]
        counts = np.where(cnt > MAX_CANDIDATES, OVERFLOW, cnt).astype(np.uint8)
        return cls(float(x0), float(y0), float(cell), int(dim),
                   torch.from_numpy(cand).to(dev), torch.from_numpy(counts).to(dev))

    @property
    def num_cells(self) -> int:
        return self.dim * self.dim

    def struct(self) -> qpip_grid:
        return qpip_grid(self.x0, self.y0, self.cell, self.dim, self.candidates.data_ptr(),
                         self.counts.data_ptr())

    def cells_of(self, xy: torch.Tensor) -> torch.Tensor:
        """Grid cell of each interleaved point (the kernels' function)."""
        p = xy.view(-1, 2)
        ix = _grid_coord_t(p[:, 0], self.x0, self.cell, self.dim)
        iy = _grid_coord_t(p[:, 1], self.y0, self.cell, self.dim)
        return iy * self.dim + ix

    def slots_of(self, xy: torch.Tensor, polygon: torch.Tensor) -> torch.Tensor:
        """Slot of `polygon` in each point's candidate list, -1 when the point has
        no polygon, its cell overflowed, or the polygon is not listed."""
        c = self.cells_of(xy)
        rows = self.candidates.view(-1, MAX_CANDIDATES)[c]
        hit = rows == polygon.to(torch.int32)[:, None]
        slot = torch.where(hit.any(1), hit.to(torch.int8).argmax(1).to(torch.int32),
                           torch.full_like(polygon, -1, dtype=torch.int32))
        ok = (polygon >= 0) & (self.counts[c].to(torch.int32) <= MAX_CANDIDATES)
        return torch.where(ok, slot, torch.full_like(slot, -1))


def random_permutation(n: int, gen: torch.Generator, device) -> torch.Tensor:
    """A random permutation of 0..n-1 (int64): random 32-bit keys, stably sorted
    with the library's radix sort (ties keep index order)."""
    keys = torch.randint(-(1 << 31), 1 << 31, (n,), generator=gen, device=device,
                         dtype=torch.int64).to(torch.int32)
    idx = torch.arange(n, dtype=torch.int32, device=device)
    _capi.sort_pairs(keys, idx, 32)
    return idx.to(torch.int64)


class Ranker:
    """MLP parameters (fp32 master copy, bf16 copy for the kernels) and Adam state."""

    def __init__(self, num_cells: int, device=None, seed: int = 0, zero: bool = False):
        L = _capi.lib()
        n = int(L.qpip_ranker_num_params(num_cells))
        self.num_cells = num_cells
        dev = torch.device(device) if device is not None else torch.device("cuda")
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        p = torch.zeros(n, dtype=torch.float32)
        if not zero:
            o = self.offsets()
            p[o["e"]:o["w1"]] = torch.randn(o["w1"] - o["e"], generator=g) * 0.5
            p[o["w1"]:o["b1"]] = torch.randn(HID * IN, generator=g) * (2.0 / IN) ** 0.5
            p[o["w2"]:o["b2"]] = torch.randn(HID * HID, generator=g) * (2.0 / HID) ** 0.5
            p[o["w3"]:o["b3"]] = torch.randn(MAX_CANDIDATES * HID, generator=g) * (1.0 / HID) ** 0.5
        self.params = p.to(dev)
        self.params_bf16 = self.params.to(torch.bfloat16)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.stats = torch.zeros(3, dtype=torch.float32, device=dev)
        self.step = 0

    def offsets(self) -> dict:
        c = self.num_cells
        o = {"e": 0, "w1": c * EMB}
        o["b1"] = o["w1"] + HID * IN
        o["w2"] = o["b1"] + HID
        o["b2"] = o["w2"] + HID * HID
        o["w3"] = o["b2"] + HID
        o["b3"] = o["w3"] + MAX_CANDIDATES * HID
        o["total"] = o["b3"] + MAX_CANDIDATES
        return o

    def struct(self) -> qpip_ranker:
        return qpip_ranker(self.params.data_ptr(), self.params_bf16.data_ptr(), self.num_cells)

    def set_params(self, p: torch.Tensor) -> None:
        self.params.copy_(p.to(self.params.device, torch.float32))
        self.params_bf16.copy_(self.params.to(torch.bfloat16))

    # -- kernels ---------------------------------------------------------------
    def features(self, grid: CandidateGrid, xy: torch.Tensor):
        n = xy.numel() // 2
        X = torch.empty((n, IN), dtype=torch.int16, device=xy.device)
        cells = torch.empty(n, dtype=torch.int32, device=xy.device)
        r, g = self.struct(), grid.struct()
        _capi.check(_capi.lib().qpip_ranker_features(ctypes.addressof(r), ctypes.addressof(g),
                                                     xy.data_ptr(), n, X.data_ptr(),
                                                     cells.data_ptr(),
                                                     _capi.stream_handle(xy.device)))
        return X, cells

    def logits(self, grid: CandidateGrid, xy: torch.Tensor) -> torch.Tensor:
        n = xy.numel() // 2
        out = torch.empty((n, MAX_CANDIDATES), dtype=torch.float32, device=xy.device)
        r, g = self.struct(), grid.struct()
        _capi.check(_capi.lib().qpip_ranker_logits(ctypes.addressof(r), ctypes.addressof(g),
                                                   xy.data_ptr(), n, out.data_ptr(),
                                                   _capi.stream_handle(xy.device)))
        return out

    def gradients(self, grid: CandidateGrid, xy: torch.Tensor, labels: torch.Tensor):
        """(grad, stats) for one batch; stats = (loss sum, correct, labelled)."""
        n = xy.numel() // 2
        labels = labels.to(torch.int32).contiguous()
        self.stats.zero_()
        r, g = self.struct(), grid.struct()
        _capi.check(_capi.lib().qpip_ranker_gradients(
            ctypes.addressof(r), ctypes.addressof(g), xy.data_ptr(), labels.data_ptr(), n,
            self.grad.data_ptr(), self.stats.data_ptr(), _capi.stream_handle(xy.device)))
        return self.grad, self.stats

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        _capi.check(_capi.lib().qpip_ranker_adam(
            self.params.data_ptr(), self.params_bf16.data_ptr(), self.grad.data_ptr(),
            self.m.data_ptr(), self.v.data_ptr(), self.params.numel(), lr, beta1, beta2, eps,
            self.step, _capi.stream_handle(self.params.device)))

    def train(self, grid: CandidateGrid, xy: torch.Tensor, labels: torch.Tensor,
              epochs: int = 3, batch: int = 16384, lr: float = 3e-3, seed: int = 0,
              log=None) -> list:
        """Adam on shuffled mini-batches. Returns per-epoch (mean loss, top-1 accuracy)."""
        pts = xy.view(-1, 2)
        n = pts.shape[0]
        gen = torch.Generator(device=pts.device)
        gen.manual_seed(seed)
        history = []
        for ep in range(epochs):
            perm = random_permutation(n, gen, pts.device)
            tot = torch.zeros(3, dtype=torch.float64, device=pts.device)
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                bxy = pts[idx].contiguous().view(-1)
                _, st = self.gradients(grid, bxy, labels[idx])
                tot += st.to(torch.float64)
                self.adam(lr)
            t = tot.cpu().numpy()
            history.append((t[0] / max(t[2], 1), t[1] / max(t[2], 1)))
            if log:
                log(f"epoch {ep}: loss {history[-1][0]:.4f} top-1 {history[-1][1]:.4f}")
        return history


@dataclass
class LocateResult:
    polygon: torch.Tensor  # int32 per point: polygon, -1 none
    tests: torch.Tensor    # uint8 per point: exact tests run by the learned pass
    flags: torch.Tensor    # uint8: bit 0 quadtree path, bit 1 first prediction not the answer
    fallback: int          # points that took the quadtree path
    extra_pairs: tuple     # (polygon, point) pairs beyond one per point, from the fallback
    seconds: dict

    def pairs(self):
        """(polygon_index, point_index) sorted by point, then polygon (int64)."""
        pt, poly = _capi.compact_nonnegative(self.polygon)
        if not self.extra_pairs[0].numel():
            return poly.to(torch.int64), pt
        # A point with extra pairs has its smallest polygon in self.polygon and the
        # rest, ascending, in extra_pairs: a stable sort by point keeps that order.
        return sort_by_point(torch.cat([poly.to(torch.int64), self.extra_pairs[0]]),
                             torch.cat([pt, self.extra_pairs[1]]), self.polygon.numel())


def sort_by_point(poly: torch.Tensor, pt: torch.Tensor, n: int):
    """(poly, pt) reordered by point with the library's stable radix sort (pairs
    of one point keep their order); n bounds the point ids. int64 in and out."""
    if pt.numel() == 0:
        return poly.to(torch.int64), pt.to(torch.int64)
    keys = pt.to(torch.int32).contiguous()
    vals = poly.to(torch.int32).contiguous()
    _capi.sort_pairs(keys, vals, max(1, int(max(n, 1) - 1).bit_length()))
    return vals.to(torch.int64) & 0xFFFFFFFF, keys.to(torch.int64) & 0xFFFFFFFF


def join_pairs_by_point(pi: torch.Tensor, hits, n: int):
    """The quadtree PIP hits as (polygon_index, original point index), sorted by
    point, then polygon: the PIP emits a point's hits in pair order, which is
    ascending polygon within its leaf, so a stable sort by point suffices."""
    pos = hits["point_index"].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    pt = pi.view(torch.int32)[pos]
    return sort_by_point(hits["polygon_index"].view(torch.int32), pt, n)


def plain_pairs(points: Points, polygons: Polygons, bbox, scale, max_depth, max_size):
    """The plain quadtree join as (polygon_index, original point index) pairs,
    sorted by point then polygon."""
    x0, x1, y0, y1 = bbox
    pi, tree = quadtree_on_points(points, x0, x1, y0, y1, scale, max_depth, max_size)
    bb = polygon_bounding_boxes(polygons)
    pairs = join_quadtree_and_bounding_boxes(tree, bb, x0, x1, y0, y1, scale, max_depth)
    hits = quadtree_point_in_polygon(pairs, tree, pi, points, polygons)
    return join_pairs_by_point(pi, hits, len(points))


def per_point_polygon(poly: torch.Tensor, pt: torch.Tensor, n: int) -> torch.Tensor:
    """One polygon per point (-1 none) from pairs; raises if a point has two."""
    out = torch.full((n,), -1, dtype=torch.int32, device=poly.device)
    if pt.numel() and bool((pt[1:] == pt[:-1]).any()):
        raise ValueError("a point lies in more than one polygon (not a partition)")
    out[pt] = poly.to(torch.int32)
    return out


def _polygon_edges(polygons: Polygons):
    """Every ring edge (v_k, v_{k+1 mod n}) of every polygon, zero-length ones
    dropped: (edge polygon id, ax, ay, bx, by) sorted by polygon, plus each
    polygon's first edge and edge count."""
    part = polygons.part_offset.to(torch.int64)
    ring = polygons.ring_offset.to(torch.int64)
    dev = ring.device
    nr = ring.numel() - 1
    rlen = ring[1:] - ring[:-1]
    v = torch.arange(int(ring[-1]) if nr > 0 else 0, device=dev)
    r_of_v = torch.repeat_interleave(torch.arange(nr, device=dev), rlen)
    nxt = v + 1
    nxt = torch.where(nxt == ring[1:][r_of_v], ring[:-1][r_of_v], nxt)
    p_of_r = torch.repeat_interleave(torch.arange(part.numel() - 1, device=dev), part[1:] - part[:-1])
    ax, ay = polygons.x[v], polygons.y[v]
    bx, by = polygons.x[nxt], polygons.y[nxt]
    keep = (ax != bx) | (ay != by)
    pid = p_of_r[r_of_v][keep]
    cnt = torch.bincount(pid, minlength=part.numel() - 1)
    first = torch.cumsum(cnt, 0) - cnt
    return pid, ax[keep], ay[keep], bx[keep], by[keep], first, cnt


def _pair_rows(first_a, cnt_a, first_b, cnt_b):
    """For pairs (a, b): every (row of a's range, row of b's range) combination."""
    k = cnt_a * cnt_b
    tot = int(k.sum())
    pair = torch.repeat_interleave(torch.arange(k.numel(), device=k.device), k)
    local = torch.arange(tot, device=k.device) - (torch.cumsum(k, 0) - k)[pair]
    return pair, first_a[pair] + local // cnt_b[pair], first_b[pair] + local % cnt_b[pair]


def overlapping_polygons(polygons: Polygons, grid: CandidateGrid, rel_tol: float = 2.0**-46,
                         first_only: bool = False):
    """Pairs of polygons listed together in a grid cell whose interiors overlap:
    a vertex of one strictly inside the other (crossing parity over all rings,
    off every edge), or two edges that cross properly. Shared vertices and
    shared edges, as the cells of a partition have, do not count: their
    orientation products are exactly zero. Orientations within rel_tol of the
    largest squared coordinate count as zero (the rounding level of the
    products), so an overlap thinner than that is not reported. Two polygons
    whose interiors overlap have overlapping boxes, and so share a listed cell
    unless the overlap lies only in overflowed cells, whose points take the
    quadtree path. Returns (i, j) int64 tensors (empty for a partition); with
    first_only, the pairs of the first batch that has any."""
    dev = grid.candidates.device
    cand = grid.candidates.view(-1, MAX_CANDIDATES).to(torch.int64)
    listed = grid.counts.to(torch.int64) <= MAX_CANDIDATES
    a, b = torch.triu_indices(MAX_CANDIDATES, MAX_CANDIDATES, 1, device=dev)
    ci, cj = cand[:, a], cand[:, b]
    ok = (ci >= 0) & (cj >= 0) & listed[:, None]
    P = polygons.part_offset.numel() - 1
    key = torch.unique(torch.minimum(ci, cj)[ok] * P + torch.maximum(ci, cj)[ok])
    pi, pj = key // P, key % P
    if pi.numel() == 0:
        return pi, pj
    pid, ax, ay, bx, by, first, cnt = _polygon_edges(polygons)
    scale = float(torch.max(torch.cat([ax.abs(), ay.abs()])).item()) if ax.numel() else 1.0
    tol = rel_tol * max(scale, 1.0) ** 2
    hit_i, hit_j = [], []
    # both directions of the vertex test; edges a -> b of one polygon are its
    # vertices (each vertex starts one edge)
    chunk = 1 << 22
    pairs_i = torch.cat([pi, pj])
    pairs_j = torch.cat([pj, pi])
    work = cnt[pairs_i] * cnt[pairs_j]
    bounds = torch.cumsum(work, 0)
    start = 0
    n = pairs_i.numel()
    while start < n:
        base = int(bounds[start - 1]) if start else 0
        end = int(torch.searchsorted(bounds, torch.tensor(base + chunk, device=dev)).item())
        end = max(end, start + 1)
        end = min(end, n)
        I, J = pairs_i[start:end], pairs_j[start:end]
        pr, ei, ej = _pair_rows(first[I], cnt[I], first[J], cnt[J])
        # vertex (ax[ei], ay[ei]) of polygon I against edge ej of polygon J
        px, py = ax[ei], ay[ei]
        qax, qay, qbx, qby = ax[ej], ay[ej], bx[ej], by[ej]
        o = (qbx - qax) * (py - qay) - (qby - qay) * (px - qax)
        on = (o.abs() <= tol) & (px >= torch.minimum(qax, qbx)) & (px <= torch.maximum(qax, qbx)) \
            & (py >= torch.minimum(qay, qby)) & (py <= torch.maximum(qay, qby))
        straddle = (qay > py) != (qby > py)
        xint = qax + (qbx - qax) * (py - qay) / torch.where(qby != qay, qby - qay,
                                                            torch.ones_like(qay))
        cross = straddle & (px < xint)
        nv = ax.numel()
        # per (pair, vertex): crossing parity and any on-edge (a vertex is the
        # start of one edge, so ei names it)
        vk = pr * nv + ei
        uk, inv = torch.unique(vk, return_inverse=True)
        parity = torch.zeros(uk.numel(), dtype=torch.int64, device=dev).scatter_add_(
            0, inv, cross.to(torch.int64)) & 1
        onany = torch.zeros(uk.numel(), dtype=torch.int64, device=dev).scatter_add_(
            0, inv, on.to(torch.int64)) > 0
        inside = (parity == 1) & ~onany
        # proper crossings of edge ei (of I) with edge ej (of J); each unordered
        # pair appears in both directions, so one direction suffices
        fwd = I[pr] < J[pr]
        o1 = (qbx - qax) * (ay[ei] - qay) - (qby - qay) * (ax[ei] - qax)
        o2 = (qbx - qax) * (by[ei] - qay) - (qby - qay) * (bx[ei] - qax)
        o3 = (bx[ei] - ax[ei]) * (qay - ay[ei]) - (by[ei] - ay[ei]) * (qax - ax[ei])
        o4 = (bx[ei] - ax[ei]) * (qby - ay[ei]) - (by[ei] - ay[ei]) * (qbx - ax[ei])
        proper = fwd & (((o1 > tol) & (o2 < -tol)) | ((o1 < -tol) & (o2 > tol))) & \
            (((o3 > tol) & (o4 < -tol)) | ((o3 < -tol) & (o4 > tol)))
        bad_pairs = torch.cat([torch.unique(uk[inside] // nv), torch.unique(pr[proper])])
        if bad_pairs.numel():
            bp = torch.unique(bad_pairs)
            hit_i.append(torch.minimum(I[bp], J[bp]))
            hit_j.append(torch.maximum(I[bp], J[bp]))
            if first_only:
                break
        start = end
    if not hit_i:
        z = torch.zeros(0, dtype=torch.int64, device=dev)
        return z, z
    k = torch.unique(torch.cat(hit_i) * P + torch.cat(hit_j))
    return k // P, k % P


class LearnedLocator:
    """Grid + ranker + exact test, with the plain quadtree path as the fallback.

    The locator answers each point with the first candidate whose exact test
    holds, which is the plain join's answer only when no two polygons overlap.
    Construction checks that (``overlapping_polygons``) and raises ValueError for
    a polygon set that is not a partition."""

    def __init__(self, polygons: Polygons, grid: CandidateGrid, ranker: Ranker, bbox, scale,
                 max_depth, max_size, check_partition: bool = True):
        if check_partition:
            oi, oj = overlapping_polygons(polygons, grid, first_only=True)
            if oi.numel():
                raise ValueError(
                    f"the learned locator needs non-overlapping polygons: {oi.numel()} pairs "
                    f"overlap (e.g. polygons {int(oi[0])} and {int(oj[0])}); use the plain "
                    "quadtree join for this polygon set")
        self.polygons = polygons
        self.grid = grid
        self.ranker = ranker
        self.bbox = bbox
        self.scale = scale
        self.max_depth = max_depth
        self.max_size = max_size

    def save(self, path, meta: dict | None = None) -> None:
        """Writes the whole index to one versioned file (io.save_index): the
        polygons, the candidate grid, the ranker's parameters and the quadtree
        parameters of the fallback. ``load`` restores a locator that answers
        every point exactly as this one does."""
        from . import io

        P, g = self.polygons, self.grid
        arrays = {"part_offset": P.part_offset, "ring_offset": P.ring_offset, "px": P.x,
                  "py": P.y, "grid_candidates": g.candidates, "grid_counts": g.counts,
                  "ranker_params": self.ranker.params}
        info = {"kind": "learned_locator", "grid": {"x0": g.x0, "y0": g.y0, "cell": g.cell,
                                                     "dim": g.dim},
                "num_cells": self.ranker.num_cells, "bbox": [float(v) for v in self.bbox],
                "scale": float(self.scale), "max_depth": int(self.max_depth),
                "max_size": int(self.max_size), "ranker_step": self.ranker.step,
                "meta": meta or {}}
        io.save_index(path, {k: v.detach().cpu().numpy() for k, v in arrays.items()}, info)

    @classmethod
    def load(cls, path, device=None) -> "LearnedLocator":
        """The locator saved by ``save``, on `device` (the current GPU by
        default). The partition was checked when the saved locator was built."""
        from . import io

        dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        arrays, info = io.load_index(path, device=dev)
        if info.get("kind") != "learned_locator":
            raise ValueError(f"{path}: not a saved learned locator")
        P = Polygons(arrays["part_offset"], arrays["ring_offset"], arrays["px"], arrays["py"],
                     device=dev)
        gi = info["grid"]
        grid = CandidateGrid(float(gi["x0"]), float(gi["y0"]), float(gi["cell"]), int(gi["dim"]),
                             arrays["grid_candidates"].to(torch.int32).contiguous(),
                             arrays["grid_counts"].to(torch.uint8).contiguous())
        model = Ranker(int(info["num_cells"]), device=dev, zero=True)
        if arrays["ranker_params"].numel() != model.params.numel():
            raise ValueError(f"{path}: ranker parameters do not match the grid")
        model.set_params(arrays["ranker_params"])
        model.step = int(info.get("ranker_step", 0))
        return cls(P, grid, model, tuple(info["bbox"]), info["scale"], info["max_depth"],
                   info["max_size"], check_partition=False)

    def locate(self, points: Points, max_tests: int = MAX_CANDIDATES) -> LocateResult:
        if points.xy is None:
            points = Points.from_points_xy(torch.stack([points.x, points.y], 1).reshape(-1))
        xy = points.xy
        n = len(points)
        dev = xy.device
        out = torch.empty(n, dtype=torch.int32, device=dev)
        tests = torch.empty(n, dtype=torch.uint8, device=dev)
        flags = torch.empty(n, dtype=torch.uint8, device=dev)
        P = self.polygons
        r, g = self.ranker.struct(), self.grid.struct()
        t0 = time.perf_counter()
        _capi.check(_capi.lib().qpip_ranker_locate(
            ctypes.addressof(r), ctypes.addressof(g), P.part_offset.data_ptr(), P.part_offset.numel(),
            P.ring_offset.data_ptr(), P.ring_offset.numel(), P.x.data_ptr(), P.y.data_ptr(),
            P.x.numel(), xy.data_ptr(), n, int(max_tests), out.data_ptr(), tests.data_ptr(),
            flags.data_ptr(), _capi.stream_handle(dev)))
        t1 = time.perf_counter()
        fb, _ = _capi.compact_nonnegative(
            torch.where(out == -2, 0, -1).to(torch.int32), with_values=False)
        extra = (torch.zeros(0, dtype=torch.int64, device=dev),
                 torch.zeros(0, dtype=torch.int64, device=dev))
        if fb.numel():
            sub = Points.from_points_xy(xy.view(-1, 2)[fb].reshape(-1))
            poly, spt = plain_pairs(sub, P, self.bbox, self.scale, self.max_depth, self.max_size)
            orig = fb[spt]
            out[fb] = -1
            first = torch.ones_like(orig, dtype=torch.bool)
            if orig.numel() > 1:
                first[1:] = orig[1:] != orig[:-1]
            out[orig[first]] = poly[first].to(torch.int32)
            extra = (poly[~first], orig[~first])
        torch.cuda.current_stream(dev).synchronize()
        t2 = time.perf_counter()
        return LocateResult(out, tests, flags, int(fb.numel()), extra,
                            {"learned": t1 - t0, "fallback": t2 - t1})
