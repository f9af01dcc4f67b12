"""contains_properly for polygons against points: the reference's callers of the
point-in-polygon entry points, rebuilt on this package's GPU calls.

Mirrors
  contains_properly (three modes)   python/cuspatial/cuspatial/core/binpreds/contains.py:19-185
  the mode choice                   python/cuspatial/cuspatial/core/binpreds/feature_contains_properly.py:67-91
  GeoSeries.contains_properly for a point right-hand side, allpairs or not
      feature_contains_properly.py:93-132,
      contains_geometry_processor.py:73-159 (part -> polygon index, dedup, relabel),
      contains_geometry_processor.py:252-275 (point result)

Result tables use the reference's column names and order. cudf's merge does not
define a row order, so the all-pairs result here is sorted by
(polygon_index, point_index).
"""

from __future__ import annotations

from math import ceil, sqrt

import torch

from .spatial import (
    Points,
    Polygons,
    Table,
    join_quadtree_and_bounding_boxes,
    pairwise_point_in_polygon,
    point_in_polygon_bitmask,
    polygon_bounding_boxes,
    quadtree_on_points,
    quadtree_point_in_polygon,
)

_BRUTE_FORCE_MAX_POLYGONS = 31


def _i64(t: torch.Tensor) -> torch.Tensor:
    """A UINT32 column as int64 values (torch cannot index with uint32 on the GPU)."""
    if t.dtype == torch.uint32:
        return t.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    return t.to(torch.int64)


def _quadtree_contains_properly(points: Points, polygons: Polygons) -> Table:
    """contains.py:19-74: the quadtree over the polygons' extent at depth 15, with
    max_size = ceil(sqrt(#points)) and the minimum scale (so no clamp warning);
    hits come back with point_index remapped through point_indices and
    polygon_index renamed part_index."""
    max_depth = 15
    min_size = ceil(sqrt(len(points)))
    if len(polygons) == 0:
        return Table()
    x_max = float(polygons.x.max())
    x_min = float(polygons.x.min())
    y_max = float(polygons.y.max())
    y_min = float(polygons.y.min())
    scale = max(x_max - x_min, y_max - y_min) / ((1 << max_depth) + 2)
    point_indices, quadtree = quadtree_on_points(points, x_min, x_max, y_min, y_max, scale,
                                                 max_depth, min_size)
    poly_bboxes = polygon_bounding_boxes(polygons)
    intersections = join_quadtree_and_bounding_boxes(quadtree, poly_bboxes, x_min, x_max, y_min,
                                                     y_max, scale, max_depth)
    hits = quadtree_point_in_polygon(intersections, quadtree, point_indices, points, polygons)
    out = Table()
    out["point_index"] = point_indices.view(torch.int32)[_i64(hits["point_index"])].view(
        point_indices.dtype)
    out["part_index"] = hits["polygon_index"]
    return out


def _brute_force_contains_properly(points: Points, polygons: Polygons) -> Table:
    """contains.py:77-120 plus the stack in contains.py:177-185: every (point,
    polygon) pair whose bit is set, in point-major order, as
    (point_index, part_index, result)."""
    n, p = len(points), len(polygons)
    out = Table()
    if p == 0 or n == 0:
        out["point_index"] = torch.zeros(0, dtype=torch.int64, device=points.device)
        out["part_index"] = torch.zeros(0, dtype=torch.int64, device=points.device)
        out["result"] = torch.zeros(0, dtype=torch.bool, device=points.device)
        return out
    bits = point_in_polygon_bitmask(points, polygons).to(torch.int64)
    cols = torch.arange(p, device=bits.device, dtype=torch.int64)
    table = ((bits[:, None] >> cols[None, :]) & 1).to(torch.bool)  # [point][polygon]
    pt, part = torch.nonzero(table, as_tuple=True)  # row-major: point, then polygon
    out["point_index"] = pt
    out["part_index"] = part
    out["result"] = torch.ones(pt.numel(), dtype=torch.bool, device=bits.device)
    return out


def _pairwise_contains_properly(points: Points, polygons: Polygons) -> Table:
    """contains.py:123-169: point i against polygon i; the rows that hold."""
    res = pairwise_point_in_polygon(points, polygons)
    trues = torch.nonzero(res, as_tuple=True)[0]
    out = Table()
    out["pairwise_index"] = trues
    out["point_index"] = trues.clone()
    out["result"] = torch.ones(trues.numel(), dtype=torch.bool, device=res.device)
    return out


def contains_properly(polygons: Polygons, points: Points, mode: str = "pairwise") -> Table:
    """contains.py:172-185. mode "quadtree" -> (point_index, part_index);
    "pairwise" -> (pairwise_index, point_index, result); anything else is the
    brute-force bitmask -> (point_index, part_index, result)."""
    if mode == "quadtree":
        return _quadtree_contains_properly(points, polygons)
    elif mode == "pairwise":
        return _pairwise_contains_properly(points, polygons)
    return _brute_force_contains_properly(points, polygons)


def has_multipolygons(polygons: Polygons) -> bool:
    """True when some geometry holds more than one polygon part."""
    g = polygons.geometry_offset
    if g is None or g.numel() < 2:
        return False
    return bool(((g[1:] - g[:-1]) > 1).any())


def num_geometries(polygons: Polygons) -> int:
    """Rows of the polygon GeoSeries (geometries; a multipolygon is one row)."""
    g = polygons.geometry_offset
    return len(polygons) if g is None else max(int(g.numel()) - 1, 0)


def pip_mode(polygons: Polygons, allpairs: bool = False) -> str:
    """feature_contains_properly.py:67-91: brute force for at most 31 geometries,
    quadtree for all-pairs or multipolygons, else pairwise."""
    if num_geometries(polygons) <= _BRUTE_FORCE_MAX_POLYGONS:
        return "brute_force"
    elif allpairs or has_multipolygons(polygons):
        return "quadtree"
    return "pairwise"


def take_polygons(polygons: Polygons, index) -> Polygons:
    """polygons[index] (GeoSeries row selection) for polygons with rings."""
    idx = torch.as_tensor(index, dtype=torch.int64, device=polygons.device)
    n = len(polygons)
    if idx.numel() and (int(idx.min()) < -n or int(idx.max()) >= n):
        raise IndexError("polygon index out of range")
    idx = torch.where(idx < 0, idx + n, idx)
    part = polygons.part_offset.to(torch.int64)
    ring = polygons.ring_offset.to(torch.int64)
    nrings = part[idx + 1] - part[idx]
    new_part = torch.zeros(idx.numel() + 1, dtype=torch.int64, device=idx.device)
    new_part[1:] = torch.cumsum(nrings, 0)
    total_rings = int(new_part[-1])
    ring_ids = (torch.repeat_interleave(part[idx], nrings) +
                torch.arange(total_rings, device=idx.device) -
                torch.repeat_interleave(new_part[:-1], nrings))
    nverts = ring[ring_ids + 1] - ring[ring_ids]
    new_ring = torch.zeros(total_rings + 1, dtype=torch.int64, device=idx.device)
    new_ring[1:] = torch.cumsum(nverts, 0)
    total_verts = int(new_ring[-1])
    vert_ids = (torch.repeat_interleave(ring[ring_ids], nverts) +
                torch.arange(total_verts, device=idx.device) -
                torch.repeat_interleave(new_ring[:-1], nverts))
    return Polygons(new_part.to(torch.int32), new_ring.to(torch.int32), polygons.x[vert_ids],
                    polygons.y[vert_ids], device=polygons.device)


def _part_to_polygon(polygons: Polygons, part_index: torch.Tensor) -> torch.Tensor:
    """contains_geometry_processor.py:73-118: part index -> geometry index through
    geometry_offset (identity for a column of simple polygons)."""
    g = polygons.geometry_offset
    if g is None:
        return _i64(part_index)
    g = g.to(torch.int64)
    return torch.searchsorted(g, _i64(part_index), right=True) - 1


def polygon_contains_properly(polygons: Polygons, points: Points, allpairs: bool = False,
                              index=None):
    """GeoSeries.contains_properly(points, allpairs) for a polygon left-hand side
    and a point right-hand side (feature_contains_properly.py:93-132,
    contains_geometry_processor.py:120-159,252-275).

    allpairs=True -> Table(polygon_index, point_index) of every containment,
    deduplicated, polygon_index relabelled through ``index`` (the left
    GeoSeries' index; positions when None). allpairs=False -> one bool per point:
    with as many polygons as points, point i holds when polygon i properly
    contains it; otherwise point j is set for every polygon label j that contains
    some point (the reference's behaviour for unequal lengths)."""
    n = len(points)
    mode = pip_mode(polygons, allpairs)
    point_indices = torch.arange(n, device=points.device, dtype=torch.int64)
    lhs = polygons
    if mode == "pairwise":
        lhs = take_polygons(polygons, point_indices)
    pip = contains_properly(lhs, points, mode=mode)
    if "pairwise_index" in pip:
        part = point_indices[_i64(pip["pairwise_index"])]
        pt = _i64(pip["point_index"])
    elif pip.num_columns == 0:
        part = torch.zeros(0, dtype=torch.int64, device=points.device)
        pt = part
    else:
        part = _i64(pip["part_index"])
        pt = _i64(pip["point_index"])
    poly = _part_to_polygon(polygons, part)
    if poly.numel():
        pairs = torch.unique(torch.stack([poly, pt], 1), dim=0)  # sorted, deduplicated
        poly, pt = pairs[:, 0], pairs[:, 1]
    if index is not None:
        labels = torch.as_tensor(index, device=points.device)
        poly = labels[poly]
    if allpairs:
        out = Table()
        out["polygon_index"] = poly
        out["point_index"] = pt
        return out
    final = torch.zeros(n, dtype=torch.bool, device=points.device)
    if num_geometries(polygons) == n:
        match = poly == pt
        final[point_indices[poly[match]]] = True
    else:
        if poly.numel() and int(poly.max()) >= n:
            raise KeyError("polygon label outside the point series")
        final[poly] = True
    return final
