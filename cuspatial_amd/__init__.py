"""cuspatial_amd: an MI355X-native quadtree point-in-polygon spatial join.

Drop-in for the reference's quadtree PIP path (rapidsai/cuspatial
quadtree_on_points -> polygon_bounding_boxes -> join_quadtree_and_bounding_boxes
-> quadtree_point_in_polygon) backed by hand-written gfx950 HIP kernels behind
the C ABI in include/cuspatial_amd.h.
"""

from ._capi import CapacityError  # noqa: F401
from .spatial import (  # noqa: F401
    BatchPlan,
    BatchResult,
    Linestrings,
    Points,
    Polygons,
    Table,
    batched_quadtree_point_in_polygon,
    join_quadtree_and_bounding_boxes,
    linestring_bounding_boxes,
    pairwise_point_in_polygon,
    point_in_polygon,
    point_in_polygon_bitmask,
    polygon_bounding_boxes,
    quadtree_on_points,
    quadtree_point_in_polygon,
    quadtree_point_to_nearest_linestring,
)
from .binpreds import contains_properly, polygon_contains_properly  # noqa: F401
from .ranker import CandidateGrid, LearnedLocator, Ranker  # noqa: F401

__all__ = [
    "Points",
    "Polygons",
    "Table",
    "Linestrings",
    "linestring_bounding_boxes",
    "quadtree_point_to_nearest_linestring",
    "quadtree_on_points",
    "polygon_bounding_boxes",
    "join_quadtree_and_bounding_boxes",
    "quadtree_point_in_polygon",
    "point_in_polygon",
    "point_in_polygon_bitmask",
    "pairwise_point_in_polygon",
    "batched_quadtree_point_in_polygon",
    "BatchPlan",
    "BatchResult",
    "CapacityError",
    "contains_properly",
    "polygon_contains_properly",
    "CandidateGrid",
    "Ranker",
    "LearnedLocator",
]
