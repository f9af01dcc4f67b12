"""cuspatial_amd: an MI355X-native quadtree point-in-polygon spatial join.

Drop-in for the reference's quadtree PIP path (rapidsai/cuspatial
quadtree_on_points -> polygon_bounding_boxes -> join_quadtree_and_bounding_boxes
-> quadtree_point_in_polygon) backed by hand-written gfx950 HIP kernels behind
the C ABI in include/cuspatial_amd.h.
"""

from .spatial import (  # noqa: F401
    Points,
    Polygons,
    Table,
    join_quadtree_and_bounding_boxes,
    polygon_bounding_boxes,
    quadtree_on_points,
    quadtree_point_in_polygon,
)

__all__ = [
    "Points",
    "Polygons",
    "Table",
    "quadtree_on_points",
    "polygon_bounding_boxes",
    "join_quadtree_and_bounding_boxes",
    "quadtree_point_in_polygon",
]
