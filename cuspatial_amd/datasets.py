"""Seeded synthetic workloads for the quadtree point-in-polygon join.

All inputs are generated on the host with numpy's PCG64 from a recorded seed
and shared by the GPU run and the CPU baseline.

  uniform_stars    1e6 points uniform in [0, 1000)^2; 1000 star polygons with
                   centre uniform, radius U[2, 10], 8-32 vertices at sorted random
                   angles with radii jittered +-30%, closed rings. Quadtree bbox
                   (0,0)-(1000,1000), scale 1000/2^7, max_depth 7, max_size 125.
  gaussian_holes   1e8 points in [0, 1e5)^2: 80% from a mixture of 64 isotropic
                   Gaussians (centres uniform, sigma log-uniform in [200, 3000];
                   draws outside the domain are redrawn), 20% uniform. 50,000 star
                   polygons, vertex count log-normal (median 24, clipped to
                   [4, 1024]), 20% of them with 1-3 holes of 0.2-0.4x the outer
                   radius; radii U[2, 10] * 100 / sqrt(50) so that polygons cover
                   the same fraction of the domain as uniform_stars. Quadtree
                   scale 1e5/2^15, max_depth 15, max_size 256.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Workload:
    name: str
    x: np.ndarray
    y: np.ndarray
    part_offset: np.ndarray  # int32, polygons + 1
    ring_offset: np.ndarray  # int32, rings + 1
    px: np.ndarray
    py: np.ndarray
    x_min: float
    x_max: float
    y_min: float
    y_max: float
    scale: float
    max_depth: int
    max_size: int
    seed: int
    centers: np.ndarray = None  # (P, 2) polygon centres, where the generator has them

    @property
    def num_points(self) -> int:
        return int(self.x.shape[0])

    @property
    def num_polygons(self) -> int:
        return int(self.part_offset.size - 1)

    def describe(self) -> dict:
        return {
            "workload": self.name,
            "points": self.num_points,
            "polygons": self.num_polygons,
            "vertices": int(self.px.size),
            "max_depth": self.max_depth,
            "max_size": self.max_size,
            "scale": self.scale,
            "bbox": [self.x_min, self.y_min, self.x_max, self.y_max],
            "seed": self.seed,
        }


def _star_rings(rng, cx, cy, radius, counts, jitter=0.3):
    """Closed star-shaped rings: sorted random angles, radii jittered by +-jitter."""
    n = counts.size
    total = int(counts.sum())
    ring_id = np.repeat(np.arange(n), counts)
    ang = rng.uniform(0.0, 2.0 * math.pi, total)
    order = np.lexsort((ang, ring_id))  # sort angles within each ring
    ang = ang[order]
    r = np.repeat(radius, counts) * rng.uniform(1.0 - jitter, 1.0 + jitter, total)
    vx = np.repeat(cx, counts) + r * np.cos(ang)
    vy = np.repeat(cy, counts) + r * np.sin(ang)
    # close each ring by repeating its first vertex
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    out_counts = counts + 1
    out_starts = np.concatenate([[0], np.cumsum(out_counts)[:-1]])
    ox = np.empty(total + n)
    oy = np.empty(total + n)
    dst = np.arange(total) + np.repeat(out_starts - starts, counts)
    ox[dst] = vx
    oy[dst] = vy
    ox[out_starts + counts] = vx[starts]
    oy[out_starts + counts] = vy[starts]
    return ox, oy, out_counts


def _star_rings_sorted_spacings(rng, cx, cy, radius, counts, jitter=0.3):
    """Closed star-shaped rings whose sorted angles are drawn directly: the
    normalised partial sums of k + 1 exponential spacings are k sorted uniforms
    (no sort, so it scales to ~1e8 vertices)."""
    n = counts.size
    k1 = counts + 1
    total1 = int(k1.sum())
    ring_id1 = np.repeat(np.arange(n), k1)
    e = rng.exponential(1.0, total1)
    cs = np.cumsum(e)
    ends = np.cumsum(k1) - 1
    starts1 = ends - counts
    base = np.concatenate([[0.0], cs[ends[:-1]]])
    tot = cs[ends] - base
    frac = (cs - base[ring_id1]) / tot[ring_id1]
    keep = np.ones(total1, dtype=bool)
    keep[ends] = False  # the last partial sum is 1
    ang = 2.0 * math.pi * frac[keep]
    r = np.repeat(radius, counts) * rng.uniform(1.0 - jitter, 1.0 + jitter, ang.size)
    vx = np.repeat(cx, counts) + r * np.cos(ang)
    vy = np.repeat(cy, counts) + r * np.sin(ang)
    # close each ring: first vertex repeated at the end
    out_counts = counts + 1
    out_starts = np.concatenate([[0], np.cumsum(out_counts)[:-1]])
    starts = out_starts - np.arange(n)
    ox = np.empty(vx.size + n)
    oy = np.empty(vy.size + n)
    dst = np.arange(vx.size) + np.repeat(np.arange(n), counts)
    ox[dst] = vx
    oy[dst] = vy
    ox[out_starts + counts] = vx[starts]
    oy[out_starts + counts] = vy[starts]
    del starts1
    return ox, oy, out_counts


def uniform_stars(seed: int = 1001, num_points: int = 1_000_000, num_polygons: int = 1000) -> Workload:
    pts = np.random.Generator(np.random.PCG64([seed, 0]))
    rng = np.random.Generator(np.random.PCG64([seed, 1]))
    ext = 1000.0
    x = pts.uniform(0.0, ext, num_points)
    y = pts.uniform(0.0, ext, num_points)
    cx = rng.uniform(0.0, ext, num_polygons)
    cy = rng.uniform(0.0, ext, num_polygons)
    radius = rng.uniform(2.0, 10.0, num_polygons)
    counts = rng.integers(8, 33, num_polygons)
    vx, vy, rc = _star_rings(rng, cx, cy, radius, counts)
    ring_offset = np.concatenate([[0], np.cumsum(rc)]).astype(np.int32)
    part_offset = np.arange(num_polygons + 1, dtype=np.int32)
    return Workload("uniform_stars", x, y, part_offset, ring_offset, vx, vy, 0.0, ext, 0.0, ext,
                    ext / 2**7, 7, 125, seed)


def _gaussian_mixture(rng, n, ext, k=64):
    cx = rng.uniform(0.0, ext, k)
    cy = rng.uniform(0.0, ext, k)
    sigma = np.exp(rng.uniform(math.log(200.0), math.log(3000.0), k))
    n_mix = int(round(0.8 * n))
    comp = rng.integers(0, k, n_mix)
    x = np.empty(n)
    y = np.empty(n)
    x[:n_mix] = cx[comp] + sigma[comp] * rng.standard_normal(n_mix)
    y[:n_mix] = cy[comp] + sigma[comp] * rng.standard_normal(n_mix)
    bad = np.nonzero((x[:n_mix] < 0) | (x[:n_mix] >= ext) | (y[:n_mix] < 0) | (y[:n_mix] >= ext))[0]
    while bad.size:
        c = comp[bad]
        x[bad] = cx[c] + sigma[c] * rng.standard_normal(bad.size)
        y[bad] = cy[c] + sigma[c] * rng.standard_normal(bad.size)
        keep = (x[bad] < 0) | (x[bad] >= ext) | (y[bad] < 0) | (y[bad] >= ext)
        bad = bad[keep]
    x[n_mix:] = rng.uniform(0.0, ext, n - n_mix)
    y[n_mix:] = rng.uniform(0.0, ext, n - n_mix)
    return x, y


def _gaussian_mixture_device(seed, n, ext, device, k=64):
    """Same mixture, drawn on the GPU with torch's seeded Philox generator (the
    host draws ~1e7 values/s, too slow for 1e8 points). Deterministic for a given
    seed and device type."""
    import torch

    host = np.random.Generator(np.random.PCG64(seed + 7919))
    cx = torch.as_tensor(host.uniform(0.0, ext, k), device=device)
    cy = torch.as_tensor(host.uniform(0.0, ext, k), device=device)
    sigma = torch.as_tensor(np.exp(host.uniform(math.log(200.0), math.log(3000.0), k)),
                            device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    n_mix = int(round(0.8 * n))
    comp = torch.randint(0, k, (n_mix,), generator=gen, device=device)
    x = torch.empty(n, dtype=torch.float64, device=device)
    y = torch.empty(n, dtype=torch.float64, device=device)
    x[:n_mix] = cx[comp] + sigma[comp] * torch.randn(n_mix, generator=gen, device=device,
                                                      dtype=torch.float64)
    y[:n_mix] = cy[comp] + sigma[comp] * torch.randn(n_mix, generator=gen, device=device,
                                                      dtype=torch.float64)
    xm, ym = x[:n_mix], y[:n_mix]
    bad = torch.nonzero((xm < 0) | (xm >= ext) | (ym < 0) | (ym >= ext)).flatten()
    while bad.numel():
        c = comp[bad]
        xm[bad] = cx[c] + sigma[c] * torch.randn(bad.numel(), generator=gen, device=device,
                                                 dtype=torch.float64)
        ym[bad] = cy[c] + sigma[c] * torch.randn(bad.numel(), generator=gen, device=device,
                                                 dtype=torch.float64)
        keep = (xm[bad] < 0) | (xm[bad] >= ext) | (ym[bad] < 0) | (ym[bad] >= ext)
        bad = bad[keep]
    x[n_mix:] = torch.rand(n - n_mix, generator=gen, device=device, dtype=torch.float64) * ext
    y[n_mix:] = torch.rand(n - n_mix, generator=gen, device=device, dtype=torch.float64) * ext
    # torch.rand draws in [0, 1); the product can round up to ext: fold back into [0, ext)
    x[n_mix:] = torch.where(x[n_mix:] >= ext, torch.zeros_like(x[n_mix:]), x[n_mix:])
    y[n_mix:] = torch.where(y[n_mix:] >= ext, torch.zeros_like(y[n_mix:]), y[n_mix:])
    return x, y


def gaussian_holes(seed: int = 1002, num_points: int = 100_000_000,
                   num_polygons: int = 50_000, device=None) -> Workload:
    """With `device` (a torch device) the points are drawn there as torch tensors;
    otherwise on the host as numpy arrays (the polygons are always host arrays)."""
    # independent streams: the polygons do not depend on the point count or on
    # where the points are drawn
    rng = np.random.Generator(np.random.PCG64([seed, 1]))
    ext = 1e5
    if device is not None:
        x, y = _gaussian_mixture_device(seed, num_points, ext, device)
    else:
        x, y = _gaussian_mixture(np.random.Generator(np.random.PCG64([seed, 0])), num_points, ext)
    cx = rng.uniform(0.0, ext, num_polygons)
    cy = rng.uniform(0.0, ext, num_polygons)
    rscale = 100.0 / math.sqrt(50.0)
    radius = rng.uniform(2.0, 10.0, num_polygons) * rscale
    counts = np.clip(np.round(np.exp(rng.normal(math.log(24.0), 0.6, num_polygons))), 4, 1024)
    counts = counts.astype(np.int64)
    ovx, ovy, orc = _star_rings(rng, cx, cy, radius, counts)
    # holes: 20% of polygons get 1-3 holes of 0.2-0.4x the outer radius, placed
    # inside the outer ring's minimum radius (0.7 R) and apart from each other
    has_holes = rng.uniform(0.0, 1.0, num_polygons) < 0.2
    nholes = np.where(has_holes, rng.integers(1, 4, num_polygons), 0)
    hp = np.repeat(np.arange(num_polygons), nholes)
    hk = np.arange(hp.size) - np.repeat(np.cumsum(nholes) - nholes, nholes)
    hn = nholes[hp]
    hr = radius[hp] * rng.uniform(0.2, 0.4, hp.size) / np.maximum(hn, 1)
    hoff = np.where(hn > 1, 0.35 * radius[hp], 0.0)
    hang = 2.0 * math.pi * hk / np.maximum(hn, 1)
    hcx = cx[hp] + hoff * np.cos(hang)
    hcy = cy[hp] + hoff * np.sin(hang)
    hcount = np.maximum(4, counts[hp] // 2)
    hvx, hvy, hrc = _star_rings(rng, hcx, hcy, hr, hcount, jitter=0.1)
    # assemble rings polygon by polygon: outer ring, then its holes
    outer_start = np.concatenate([[0], np.cumsum(orc)[:-1]])
    hole_start = np.concatenate([[0], np.cumsum(hrc)[:-1]]) if hrc.size else np.zeros(0, np.int64)
    rings_per_poly = 1 + nholes
    xs, ys, sizes = [], [], []
    h = 0
    for i in range(num_polygons):
        a, c = outer_start[i], orc[i]
        xs.append(ovx[a:a + c])
        ys.append(ovy[a:a + c])
        sizes.append(c)
        for _ in range(nholes[i]):
            a, c = hole_start[h], hrc[h]
            xs.append(hvx[a:a + c])
            ys.append(hvy[a:a + c])
            sizes.append(c)
            h += 1
    px = np.concatenate(xs)
    py = np.concatenate(ys)
    ring_offset = np.concatenate([[0], np.cumsum(sizes)])
    part_offset = np.concatenate([[0], np.cumsum(rings_per_poly)]).astype(np.int32)
    return Workload("gaussian_holes", x, y, part_offset, ring_offset.astype(np.int32), px, py,
                    0.0, ext, 0.0, ext, ext / 2**15, 15, 256, seed, np.stack([cx, cy], 1))


def subsample(w: Workload, n: int) -> Workload:
    """The first n points of a workload as host arrays (points are i.i.d., so a
    prefix is a sample)."""
    def host(a):
        return a[:n].cpu().numpy().copy() if hasattr(a, "cpu") else np.array(a[:n])

    return Workload(w.name + f"[:{n}]", host(w.x), host(w.y), w.part_offset,
                    w.ring_offset, w.px, w.py, w.x_min, w.x_max, w.y_min, w.y_max, w.scale,
                    w.max_depth, w.max_size, w.seed)


@dataclass
class BatchWorkload:
    """Many independent jobs: job j owns points [job_point_offsets[j], [j+1]) of the
    interleaved xy array and polygons [job_polygon_offsets[j], [j+1]) of the
    GeoArrow polygon arrays, with its own bbox, scale, depth and max_size."""

    name: str
    xy: object  # interleaved x, y (numpy array or torch tensor)
    job_point_offsets: np.ndarray
    part_offset: np.ndarray
    ring_offset: np.ndarray
    px: object
    py: object
    job_polygon_offsets: np.ndarray
    bboxes: np.ndarray  # (J, 4): x_min, x_max, y_min, y_max
    scales: np.ndarray
    max_depths: np.ndarray
    max_sizes: np.ndarray
    seed: int

    @property
    def num_jobs(self) -> int:
        return len(self.job_point_offsets) - 1

    @property
    def num_points(self) -> int:
        return int(self.job_point_offsets[-1])

    def job_points(self, j):
        """(x, y) of job j as host arrays."""
        a, b = int(self.job_point_offsets[j]), int(self.job_point_offsets[j + 1])
        xy = self.xy[2 * a:2 * b]
        xy = xy.cpu().numpy() if hasattr(xy, "cpu") else np.asarray(xy)
        return xy[0::2].copy(), xy[1::2].copy()

    def job_polygons(self, j):
        """Job j's polygons as standalone GeoArrow arrays (offsets rebased)."""
        a, b = int(self.job_polygon_offsets[j]), int(self.job_polygon_offsets[j + 1])
        part = np.asarray(self.part_offset[a:b + 1], dtype=np.int64)
        r0, r1 = int(part[0]), int(part[-1])
        ring = np.asarray(self.ring_offset[r0:r1 + 1], dtype=np.int64)
        v0, v1 = int(ring[0]), int(ring[-1])
        px = self.px.cpu().numpy() if hasattr(self.px, "cpu") else np.asarray(self.px)
        py = self.py.cpu().numpy() if hasattr(self.py, "cpu") else np.asarray(self.py)
        return ((part - r0).astype(np.int32), (ring - v0).astype(np.int32),
                px[v0:v1].copy(), py[v0:v1].copy())


def batched_jobs(seed: int = 1003, num_jobs: int = 4096, min_points: int = 1_000,
                 max_points: int = 1_000_000, min_polygons: int = 1, max_polygons: int = 2000,
                 min_vertices: int = 4, max_vertices: int = 256, ext: float = 1e4,
                 device=None) -> BatchWorkload:
    """The batched-join workload: per job, log-uniform point and polygon counts,
    points uniform in a job-local ext-wide square, star polygons (closed single
    rings, radius U[2,10] * ext/1000, vertex count uniform), max_depth =
    ceil(log4(points / 128)) (at least 1), max_size 128, scale = ext / 2^depth.
    With `device` the points are drawn there (torch, seeded) and returned as a
    tensor; the polygons are host arrays."""
    rng = np.random.Generator(np.random.PCG64([seed, 2]))
    npts = np.round(np.exp(rng.uniform(math.log(min_points), math.log(max_points), num_jobs)))
    npts = npts.astype(np.int64)
    npoly = np.round(np.exp(rng.uniform(math.log(min_polygons), math.log(max_polygons),
                                        num_jobs))).astype(np.int64)
    depth = np.maximum(1, np.ceil(np.log(np.maximum(npts, 1) / 128.0) / math.log(4.0)))
    depth = np.minimum(depth, 15).astype(np.int8)
    jpo = np.concatenate([[0], np.cumsum(npts)]).astype(np.int64)
    jpy = np.concatenate([[0], np.cumsum(npoly)]).astype(np.int64)
    N, P = int(jpo[-1]), int(jpy[-1])
    if device is not None:
        import torch

        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
        xy = torch.rand(2 * N, generator=gen, device=device, dtype=torch.float64) * ext
        xy = torch.where(xy >= ext, torch.zeros_like(xy), xy)
    else:
        prng = np.random.Generator(np.random.PCG64([seed, 0]))
        xy = prng.uniform(0.0, ext, 2 * N)
    grng = np.random.Generator(np.random.PCG64([seed, 1]))
    cx = grng.uniform(0.0, ext, P)
    cy = grng.uniform(0.0, ext, P)
    radius = grng.uniform(2.0, 10.0, P) * (ext / 1000.0)
    counts = grng.integers(min_vertices, max_vertices + 1, P)
    vx, vy, rc = _star_rings_sorted_spacings(grng, cx, cy, radius, counts)
    ring_offset = np.concatenate([[0], np.cumsum(rc)]).astype(np.int32)
    part_offset = np.arange(P + 1, dtype=np.int32)
    bboxes = np.tile(np.array([0.0, ext, 0.0, ext]), (num_jobs, 1))
    scales = ext / np.power(2.0, depth.astype(np.float64))
    return BatchWorkload("batched_jobs", xy, jpo, part_offset, ring_offset, vx, vy, jpy, bboxes,
                         scales, depth, np.full(num_jobs, 128, np.int32), seed)


WORKLOADS = {"uniform_stars": uniform_stars, "gaussian_holes": gaussian_holes,
             "batched_jobs": batched_jobs}


# ---------------------------------------------------------------------------
# Voronoi partition (non-overlapping polygons covering the domain)
# ---------------------------------------------------------------------------
def voronoi_cells(seed: int, num_seeds: int, ext: float, seeds: np.ndarray = None):
    """Clipped Voronoi tessellation of `num_seeds` uniform seeds in [0, ext)^2 (or
    of the given `seeds`, (n, 2) points inside the square) as (part_offset,
    ring_offset, px, py): one closed counter-clockwise ring per cell. The seeds
    are mirrored across the four sides of the square so that every original cell
    is bounded and ends on the square's sides. Neighbouring cells share their
    common vertices bit for bit (one Qhull vertex each), so the cells do not
    overlap."""
    from scipy.spatial import Voronoi

    if seeds is None:
        rng = np.random.Generator(np.random.PCG64([seed, 2]))
        s = rng.uniform(0.0, ext, (num_seeds, 2))
    else:
        s = np.asarray(seeds, np.float64)
        num_seeds = s.shape[0]
    mirrored = [s,
                np.stack([-s[:, 0], s[:, 1]], 1), np.stack([2 * ext - s[:, 0], s[:, 1]], 1),
                np.stack([s[:, 0], -s[:, 1]], 1), np.stack([s[:, 0], 2 * ext - s[:, 1]], 1)]
    vor = Voronoi(np.concatenate(mirrored))
    verts = vor.vertices
    xs, ys, sizes = [], [], []
    for i in range(num_seeds):
        reg = [v for v in vor.regions[vor.point_region[i]]]
        if -1 in reg or len(reg) < 3:
            raise RuntimeError("unbounded Voronoi cell after mirroring")
        reg = np.asarray(reg)
        ang = np.arctan2(verts[reg, 1] - s[i, 1], verts[reg, 0] - s[i, 0])
        reg = reg[np.argsort(ang, kind="stable")]
        reg = np.concatenate([reg, reg[:1]])
        xs.append(verts[reg, 0])
        ys.append(verts[reg, 1])
        sizes.append(reg.size)
    ring_offset = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    part_offset = np.arange(num_seeds + 1, dtype=np.int32)
    return part_offset, ring_offset, np.concatenate(xs), np.concatenate(ys)


def voronoi_partition(seed: int = 1004, num_points: int = 200_000_000, num_seeds: int = 20_000,
                      device=None) -> Workload:
    """Points from the 64-Gaussian mixture of gaussian_holes over [0, 1e5)^2 and
    polygons from the clipped Voronoi tessellation of `num_seeds` uniform seeds
    (convex, non-overlapping, covering the domain). The quadtree parameters are
    those of gaussian_holes (max_depth 15, max_size 256, scale 1e5/2^15)."""
    ext = 1e5
    if device is not None:
        x, y = _gaussian_mixture_device(seed, num_points, ext, device)
    else:
        x, y = _gaussian_mixture(np.random.Generator(np.random.PCG64([seed, 0])), num_points, ext)
    part, ring, px, py = voronoi_cells(seed, num_seeds, ext)
    return Workload("voronoi_partition", x, y, part, ring, px, py, 0.0, ext, 0.0, ext,
                    ext / 2**15, 15, 256, seed)


def gaussian_holes_partition(seed: int = 1002, num_points: int = 0, device=None) -> Workload:
    """The Voronoi-partition variant of gaussian_holes' polygon set: the clipped
    Voronoi cells of its 50,000 polygon centres (convex, non-overlapping,
    covering [0, 1e5)^2), with gaussian_holes' points (num_points of them) and
    quadtree parameters. The resident partition index of the serving workload."""
    gh = gaussian_holes(seed=seed, num_points=max(num_points, 1), device=device)
    part, ring, px, py = voronoi_cells(seed, 0, 1e5, seeds=gh.centers)
    x, y = (gh.x, gh.y) if num_points > 0 else (gh.x[:0], gh.y[:0])
    return Workload("gaussian_holes_partition", x, y, part, ring, px, py, gh.x_min, gh.x_max,
                    gh.y_min, gh.y_max, gh.scale, gh.max_depth, gh.max_size, seed, gh.centers)
