"""Sizes past the launch-grid limits of one-item-per-thread and one-item-per-wave
kernels: more than 2^22 rings (edge table, boxes: one wave each) and more than
2^28 points in one batched group (point_indices: one thread each).

Bar (size-independent properties): 5,000,000 unit squares with a point at each
centre give exactly the pairs (i, i), on the single-job path and the batched
path; a batched group of 2^28 + 2^20 points returns the same point_indices as
the single-job quadtree_on_points, and they are a permutation of the job's points.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cs():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd

    return cuspatial_amd


def squares(n, side):
    """n closed unit squares on a side x side lattice (one ring each), centres at
    (i + 0.5, j + 0.5)."""
    k = np.arange(n)
    cx = (k % side).astype(np.float64) + 0.5
    cy = (k // side).astype(np.float64) + 0.5
    dx = np.array([-0.4, 0.4, 0.4, -0.4, -0.4])
    dy = np.array([-0.4, -0.4, 0.4, 0.4, -0.4])
    vx = (cx[:, None] + dx[None, :]).reshape(-1)
    vy = (cy[:, None] + dy[None, :]).reshape(-1)
    ring = (5 * np.arange(n + 1)).astype(np.int32)
    part = np.arange(n + 1, dtype=np.int32)
    return part, ring, vx, vy, cx, cy


def test_five_million_rings(cs):
    import torch

    n, side = 5_000_000, 2237
    part, ring, vx, vy, cx, cy = squares(n, side)
    ext = float(side)
    print(f"[large] {n} squares", flush=True)
    P = cs.Polygons(part, ring, vx, vy, device="cuda")
    pts = cs.Points(cx, cy, device="cuda")
    pi, tree = cs.quadtree_on_points(pts, 0.0, ext, 0.0, ext, ext / 2**11, 11, 64)
    bb = cs.polygon_bounding_boxes(P)
    assert not torch.isnan(bb["minx"]).any()
    pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, 0.0, ext, 0.0, ext, ext / 2**11, 11)
    hits = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, P)
    poly = hits["polygon_index"].view(torch.int32).to(torch.int64)
    pt = pi.view(torch.int32).to(torch.int64)[hits["point_index"].view(torch.int32).to(torch.int64)]
    assert poly.numel() == n
    assert torch.equal(poly, pt)  # polygon i holds point i
    assert torch.equal(torch.sort(pt).values, torch.arange(n, device="cuda"))
    # the batched path: one job holding everything
    xy = torch.stack([pts.x, pts.y], 1).reshape(-1).contiguous()
    res = cs.batched_quadtree_point_in_polygon(
        cs.Points.from_points_xy(xy), [0, n], P, [0, n], [(0.0, ext, 0.0, ext)], [ext / 2**11],
        [11], [64])
    assert res.num_hits == n
    bpi, bh = res.job(0)
    assert torch.equal(bh["polygon_index"].view(torch.int32).cpu(),
                       hits["polygon_index"].view(torch.int32).cpu())
    assert torch.equal(bh["point_index"].view(torch.int32).cpu(),
                       hits["point_index"].view(torch.int32).cpu())


def test_batched_group_past_2_28_points(cs):
    import torch

    n = (1 << 28) + (1 << 20)
    g = torch.Generator(device="cuda").manual_seed(3)
    xy = torch.rand(2 * n, generator=g, dtype=torch.float64, device="cuda") * 1000.0
    part, ring, vx, vy, _, _ = squares(16, 4)
    P = cs.Polygons(part, ring, vx * 250.0, vy * 250.0, device="cuda")
    print(f"[large] batched group of {n} points", flush=True)
    pts = cs.Points.from_points_xy(xy)
    res = cs.batched_quadtree_point_in_polygon(pts, [0, 1000, n], P, [0, 8, 16],
                                               [(0.0, 1000.0, 0.0, 1000.0)] * 2,
                                               [1000.0 / 2**9] * 2, [9, 9], [512, 512])
    bpi, _ = res.job(1)
    sub = cs.Points.from_points_xy(xy[2000:])
    pi, tree = cs.quadtree_on_points(sub, 0.0, 1000.0, 0.0, 1000.0, 1000.0 / 2**9, 9, 512)
    assert bpi.numel() == n - 1000
    assert torch.equal(bpi.view(torch.int32), pi.view(torch.int32))
    s = torch.sort(bpi.view(torch.int32)).values
    assert torch.equal(s, torch.arange(n - 1000, dtype=torch.int32, device="cuda"))
