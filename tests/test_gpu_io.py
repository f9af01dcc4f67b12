"""A saved index, loaded back onto the GPU, answers exactly as before saving."""

import numpy as np
import pytest

from conftest import star_polygons

pytestmark = pytest.mark.gpu


def test_index_file_restart(tmp_path):
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd as cs
    from cuspatial_amd import io

    rng = np.random.default_rng(31)
    part, ring, vx, vy = star_polygons(rng, 200, 1000.0, 5.0, 40.0, holes=0.3)
    x, y = rng.uniform(0, 1000, 200000), rng.uniform(0, 1000, 200000)
    pts = cs.Points(x, y, device="cuda")
    P = cs.Polygons(part, ring, vx, vy, device="cuda")
    args = (0.0, 1000.0, 0.0, 1000.0, 1000.0 / 128, 7, 100)
    pi, tree = cs.quadtree_on_points(pts, *args)
    bb = cs.polygon_bounding_boxes(P)
    pairs = cs.join_quadtree_and_bounding_boxes(tree, bb, *args[:6])
    hits = cs.quadtree_point_in_polygon(pairs, tree, pi, pts, P)

    arrays = {"point_indices": pi, **{k: v for k, v in tree.items()},
              **{"bbox_" + k: v for k, v in bb.items()},
              "part_offset": P.part_offset, "ring_offset": P.ring_offset, "px": P.x, "py": P.y}
    path = tmp_path / "index.arrow"
    io.save_index(path, {k: v.cpu().numpy() for k, v in arrays.items()},
                  {"bbox": list(args[:4]), "scale": args[4], "max_depth": args[5],
                   "max_size": args[6]})
    got, meta = io.load_index(path, device="cuda")
    tree2 = cs.Table((k, got[k]) for k in tree.keys())
    bb2 = cs.Table((k, got["bbox_" + k]) for k in bb.keys())
    P2 = cs.Polygons(got["part_offset"], got["ring_offset"], got["px"], got["py"], device="cuda")
    pairs2 = cs.join_quadtree_and_bounding_boxes(tree2, bb2, *meta["bbox"], meta["scale"],
                                                 meta["max_depth"])
    hits2 = cs.quadtree_point_in_polygon(pairs2, tree2, got["point_indices"], pts, P2)
    for k in ("polygon_index", "point_index"):
        assert torch.equal(hits[k].cpu().to(torch.int64), hits2[k].cpu().to(torch.int64))
    assert len(hits) > 1000
