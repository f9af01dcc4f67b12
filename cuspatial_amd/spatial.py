"""Python entry points of the quadtree point-in-polygon path.

Mirrors the reference's Python API (same names, argument order, clamping,
warnings and column names):
  quadtree_on_points                python/cuspatial/cuspatial/core/spatial/indexing.py:15-199
  join_quadtree_and_bounding_boxes  python/cuspatial/cuspatial/core/spatial/join.py:105-175
  quadtree_point_in_polygon         python/cuspatial/cuspatial/core/spatial/join.py:178-262
  polygon_bounding_boxes            python/cuspatial/cuspatial/core/spatial/bounding.py:19-80
  point_in_polygon                  python/cuspatial/cuspatial/core/spatial/join.py:23-101
  pairwise_point_in_polygon         cpp/include/cuspatial/point_in_polygon.hpp:124-131

cuDF is not used: points and polygons are given as :class:`Points` /
:class:`Polygons` (GeoArrow offset + coordinate arrays, as the reference's
GeoSeries holds them), and results are :class:`Table` objects whose columns are
torch tensors on the GPU. Everything runs through libcuspatial_amd's C ABI.
"""

from __future__ import annotations

import ctypes
import warnings
from collections import OrderedDict

import numpy as np
import torch

from . import _capi

_vp = ctypes.c_void_p


# ---------------------------------------------------------------------------
# Containers
# ---------------------------------------------------------------------------
def _device(device=None) -> torch.device:
    if device is not None:
        return torch.device(device)
    if not torch.cuda.is_available():
        raise RuntimeError("cuspatial_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _as_tensor(a, dtype, device, name) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        t = a
    else:
        arr = np.asarray(a)
        if arr.dtype == np.uint32:
            arr = arr.astype(np.int64)  # torch.from_numpy lacks uint32 on some builds
        t = torch.as_tensor(arr)
    if dtype == torch.float64 and t.dtype not in (torch.float64,):
        if t.is_floating_point():
            raise TypeError(f"{name}: cuspatial_amd supports float64 coordinates only "
                            f"(got {t.dtype})")
        raise RuntimeError("cuSpatial failure: Only floating-point types are supported")
    if t.dtype != dtype:
        t = t.to(dtype)
    return t.to(device).contiguous()


class Points:
    """Point coordinates. Either the x / y columns of a point GeoSeries
    (``Points(x, y)``, the column API's layout) or its interleaved xy leaf column
    (``Points.from_points_xy``, the GeoSeries storage and the vec_2d layout of the
    header-only API, geoseries.py:237-250). Interleaved points are handed to the
    library as they are; a gathered point then costs one cache line, not two."""

    def __init__(self, x, y, device=None):
        dev = _device(device if device is not None else
                      (x.device if isinstance(x, torch.Tensor) and x.is_cuda else None))
        self._x = _as_tensor(x, torch.float64, dev, "x")
        self._y = _as_tensor(y, torch.float64, dev, "y")
        self.xy = None
        self.device = dev

    @classmethod
    def from_points_xy(cls, xy, device=None) -> "Points":
        """Interleaved [x0, y0, x1, y1, ...] (GeoSeries.from_points_xy)."""
        t = xy if isinstance(xy, torch.Tensor) else torch.as_tensor(np.asarray(xy, np.float64))
        dev = _device(device if device is not None else (t.device if t.is_cuda else None))
        t = _as_tensor(t.reshape(-1), torch.float64, dev, "xy")
        if t.numel() % 2:
            raise ValueError("interleaved xy needs an even number of values")
        self = cls.__new__(cls)
        self._x = self._y = None
        self.xy = t
        self.device = dev
        return self

    @property
    def x(self) -> torch.Tensor:
        return self._x if self.xy is None else self.xy[0::2]

    @property
    def y(self) -> torch.Tensor:
        return self._y if self.xy is None else self.xy[1::2]

    def __len__(self):
        return int(self._x.numel()) if self.xy is None else int(self.xy.numel() // 2)


class Polygons:
    """Polygons as GeoArrow offsets (the part_offset / ring_offset / x / y of a
    polygon GeoSeries, python/cuspatial/cuspatial/core/spatial/join.py:238-246)."""

    def __init__(self, part_offset, ring_offset, x, y, geometry_offset=None, device=None):
        dev = _device(device if device is not None else
                      (x.device if isinstance(x, torch.Tensor) and x.is_cuda else None))
        self.part_offset = _as_tensor(part_offset, torch.int32, dev, "part_offset")
        self.ring_offset = _as_tensor(ring_offset, torch.int32, dev, "ring_offset")
        self.x = _as_tensor(x, torch.float64, dev, "x")
        self.y = _as_tensor(y, torch.float64, dev, "y")
        self.geometry_offset = (None if geometry_offset is None else
                                _as_tensor(geometry_offset, torch.int32, dev, "geometry_offset"))
        self.device = dev

    @classmethod
    def from_polygons_xy(cls, xy, ring_offset, part_offset, geometry_offset=None,
                         device=None) -> "Polygons":
        """GeoSeries.from_polygons_xy(xy, ring_offset, part_offset, geometry_offset)."""
        t = xy if isinstance(xy, torch.Tensor) else torch.as_tensor(np.asarray(xy, np.float64))
        t = t.reshape(-1, 2)
        return cls(part_offset, ring_offset, t[:, 0].contiguous(), t[:, 1].contiguous(),
                   geometry_offset=geometry_offset, device=device)

    def __len__(self):
        return max(int(self.part_offset.numel()) - 1, 0)


class Linestrings:
    """Linestrings as GeoArrow offsets (the part_offset / x / y of a linestring
    GeoSeries' ``lines`` accessor); ``geometry_offset`` groups parts into
    multilinestrings."""

    def __init__(self, part_offset, x, y, geometry_offset=None, device=None):
        dev = _device(device if device is not None else
                      (x.device if isinstance(x, torch.Tensor) and x.is_cuda else None))
        self.part_offset = _as_tensor(part_offset, torch.int32, dev, "part_offset")
        self.x = _as_tensor(x, torch.float64, dev, "x")
        self.y = _as_tensor(y, torch.float64, dev, "y")
        self.geometry_offset = (None if geometry_offset is None else
                                _as_tensor(geometry_offset, torch.int32, dev, "geometry_offset"))
        self.device = dev

    def __len__(self):
        g = self.geometry_offset
        return max(int((g if g is not None else self.part_offset).numel()) - 1, 0)


class Table(OrderedDict):
    """An ordered set of equal-length device columns (what the reference returns
    as a cudf.DataFrame)."""

    def __len__(self):  # number of rows, like DataFrame.__len__
        for v in self.values():
            return int(v.numel())
        return 0

    @property
    def columns(self):
        return list(self.keys())

    @property
    def num_columns(self):
        return OrderedDict.__len__(self)

    def to_numpy(self) -> "OrderedDict[str, np.ndarray]":
        return OrderedDict((k, v.cpu().numpy()) for k, v in self.items())

    def to_pandas(self):
        import pandas as pd

        return pd.DataFrame(self.to_numpy())


def _p(t: torch.Tensor | None):
    return _vp(t.data_ptr()) if t is not None and t.numel() > 0 else None


def _outs(n):
    return [_vp() for _ in range(n)]


# ---------------------------------------------------------------------------
# Entry points
# ---------------------------------------------------------------------------
def quadtree_on_points(points: Points, x_min, x_max, y_min, y_max, scale, max_depth, max_size):
    """Construct a quadtree from a set of points for a given area-of-interest
    bounding box (indexing.py:15-199). Returns ``(point_indices, quadtree)``
    where quadtree has columns key, level, is_internal_node, length, offset."""
    if not isinstance(points, Points):
        raise ValueError("GeoSeries must contain only points.")
    # indexing.py:165-189: order the bbox, clamp the scale with a warning
    x_min, x_max, y_min, y_max = (min(x_min, x_max), max(x_min, x_max),
                                  min(y_min, y_max), max(y_min, y_max))
    min_scale = max(x_max - x_min, y_max - y_min) / ((1 << max_depth) + 2)
    if scale < min_scale:
        warnings.warn(
            "scale {} is less than required minimum ".format(scale)
            + "scale {}. Clamping to minimum scale".format(min_scale))
    dev = points.device
    res = _capi.TorchResource(dev)
    pi, key, level, internal, length, offset = _outs(6)
    m = ctypes.c_int64()
    L = _capi.lib()
    common = (float(x_min), float(x_max), float(y_min), float(y_max),
              float(max(scale, min_scale)), ctypes.c_int8(int(max_depth)),
              ctypes.c_int32(int(max_size)), ctypes.byref(res.struct), _capi.stream_handle(dev),
              ctypes.byref(pi), ctypes.byref(key), ctypes.byref(level), ctypes.byref(internal),
              ctypes.byref(length), ctypes.byref(offset), ctypes.byref(m))
    if points.xy is not None:
        _capi.check(L.qpip_quadtree_on_points_xy(_p(points.xy), len(points), *common))
    else:
        _capi.check(L.qpip_quadtree_on_points(_p(points.x), _p(points.y), points.x.numel(),
                                              points.y.numel(), *common))
    n = len(points)
    M = m.value
    point_indices = res.take(pi.value, n if M > 0 else 0, torch.uint32)
    tree = Table()
    tree["key"] = res.take(key.value, M, torch.uint32)
    tree["level"] = res.take(level.value, M, torch.uint8)
    tree["is_internal_node"] = res.take(internal.value, M, torch.bool)
    tree["length"] = res.take(length.value, M, torch.uint32)
    tree["offset"] = res.take(offset.value, M, torch.uint32)
    return point_indices, tree


def polygon_bounding_boxes(polygons: Polygons):
    """Minimum bounding box of each polygon (bounding.py:19-80). Columns minx,
    miny, maxx, maxy."""
    names = ["minx", "miny", "maxx", "maxy"]
    if len(polygons) == 0:
        dev = polygons.device if isinstance(polygons, Polygons) else _device()
        return Table((k, torch.empty(0, dtype=torch.float64, device=dev)) for k in names)
    if not isinstance(polygons, Polygons):
        raise ValueError("Geoseries must contain only polygons.")
    dev = polygons.device
    res = _capi.TorchResource(dev)
    outs = _outs(4)
    nb = ctypes.c_int64()
    _capi.check(_capi.lib().qpip_polygon_bounding_boxes(
        _p(polygons.part_offset), polygons.part_offset.numel(), _p(polygons.ring_offset),
        polygons.ring_offset.numel(), _p(polygons.x), _p(polygons.y), polygons.x.numel(),
        polygons.y.numel(), 0.0, ctypes.byref(res.struct), _capi.stream_handle(dev),
        *[ctypes.byref(o) for o in outs], ctypes.byref(nb)))
    return Table((k, res.take(o.value, nb.value, torch.float64)) for k, o in zip(names, outs))


def _u32(t: torch.Tensor, name) -> torch.Tensor:
    """A UINT32 column as 4-byte words (other integer widths keep their low 32
    bits, which is the value for every index that fits the column type)."""
    if t.dtype in (torch.uint32, torch.int32):
        return t.contiguous()
    if t.dtype.is_floating_point or t.dtype == torch.bool:
        raise TypeError(f"{name}: expected an integer column (got {t.dtype})")
    return t.to(torch.int32).contiguous()


def _u8(t: torch.Tensor, name) -> torch.Tensor:
    if t.dtype.is_floating_point:
        raise TypeError(f"{name}: expected an integer column (got {t.dtype})")
    if t.dtype == torch.bool:  # one byte holding 0 or 1: the BOOL8 column as it is
        return t.contiguous().view(torch.uint8)
    return t.to(torch.uint8).contiguous()


def _tree_columns(quadtree):
    cols = [quadtree[k] for k in ("key", "level", "is_internal_node", "length", "offset")]
    if len(cols) != 5:
        raise RuntimeError("cuSpatial failure: quadtree table must have 5 columns")
    key, level, internal, length, offset = cols
    return (_u32(key, "key"), _u8(level, "level"), _u8(internal, "is_internal_node"),
            _u32(length, "length"), _u32(offset, "offset"))


def join_quadtree_and_bounding_boxes(quadtree, bounding_boxes, x_min, x_max, y_min, y_max,
                                     scale, max_depth):
    """Search a quadtree for bounding box intersections (join.py:105-175).
    Returns columns bbox_offset, quad_offset."""
    x_min, x_max, y_min, y_max = (min(x_min, x_max), max(x_min, x_max),
                                  min(y_min, y_max), max(y_min, y_max))
    min_scale = max(x_max - x_min, y_max - y_min) / ((1 << max_depth) + 2)
    if scale < min_scale:
        warnings.warn(
            "scale {} is less than required minimum ".format(scale)
            + "scale {}. Clamping to minimum scale".format(min_scale))
    ncols = quadtree.num_columns if isinstance(quadtree, Table) else len(quadtree)
    if ncols != 5:
        raise RuntimeError("cuSpatial failure: quadtree table must have 5 columns")
    bcols = list(bounding_boxes.values()) if isinstance(bounding_boxes, dict) else list(bounding_boxes)
    if len(bcols) != 4:
        raise RuntimeError("cuSpatial failure: bbox table must have 4 columns")
    key, level, internal, length, offset = _tree_columns(quadtree)
    dev = key.device
    bcols = [_as_tensor(c, torch.float64, dev, "bbox") for c in bcols]
    res = _capi.TorchResource(dev)
    bo, qo = _outs(2)
    k = ctypes.c_int64()
    _capi.check(_capi.lib().qpip_join_quadtree_and_bounding_boxes(
        _p(key), _p(level), _p(internal), _p(length), _p(offset), key.numel(),
        *[_p(c) for c in bcols], bcols[0].numel(), float(x_min), float(x_max), float(y_min),
        float(y_max), float(max(scale, min_scale)), ctypes.c_int8(int(max_depth)),
        ctypes.byref(res.struct), _capi.stream_handle(dev), ctypes.byref(bo), ctypes.byref(qo),
        ctypes.byref(k)))
    out = Table()
    out["bbox_offset"] = res.take(bo.value, k.value, torch.uint32)
    out["quad_offset"] = res.take(qo.value, k.value, torch.uint32)
    return out


def quadtree_point_in_polygon(poly_quad_pairs, quadtree, point_indices, points: Points,
                              polygons: Polygons):
    """Test whether the specified points are inside any of the specified
    polygons (join.py:178-262). Returns columns polygon_index, point_index;
    point_index is a position in ``point_indices``."""
    if not isinstance(points, Points):
        raise ValueError("`point` Geoseries must contains only point geometries.")
    if not isinstance(polygons, Polygons):
        raise ValueError("`polygons` Geoseries must contains only polygons geometries.")
    pcols = list(poly_quad_pairs.values()) if isinstance(poly_quad_pairs, dict) else list(poly_quad_pairs)
    if len(pcols) != 2:
        raise RuntimeError("cuSpatial failure: a quadrant-polygon table must have 2 columns")
    key, level, internal, length, offset = _tree_columns(quadtree)
    dev = points.device
    pp = _u32(pcols[0], "poly_quad_pairs")
    pq = _u32(pcols[1], "poly_quad_pairs")
    pi = _u32(point_indices, "point_indices")
    res = _capi.TorchResource(dev)
    po, pt = _outs(2)
    h = ctypes.c_int64()
    head = (_p(pp), _p(pq), pp.numel(), _p(key), _p(level), _p(internal), _p(length),
            _p(offset), key.numel(), _p(pi), pi.numel())
    tail = (_p(polygons.part_offset), polygons.part_offset.numel(), _p(polygons.ring_offset),
            polygons.ring_offset.numel(), _p(polygons.x), _p(polygons.y), polygons.x.numel(),
            polygons.y.numel(), ctypes.byref(res.struct), _capi.stream_handle(dev),
            ctypes.byref(po), ctypes.byref(pt), ctypes.byref(h))
    L = _capi.lib()
    if points.xy is not None:
        _capi.check(L.qpip_quadtree_point_in_polygon_xy(*head, _p(points.xy), len(points), *tail))
    else:
        _capi.check(L.qpip_quadtree_point_in_polygon(*head, _p(points.x), _p(points.y),
                                                     points.x.numel(), points.y.numel(), *tail))
    out = Table()
    out["polygon_index"] = res.take(po.value, h.value, torch.uint32)
    out["point_index"] = res.take(pt.value, h.value, torch.uint32)
    return out


def _brute_force(points: Points, polygons: Polygons, pairwise: bool):
    if not isinstance(points, Points):
        raise ValueError("`points` must contain only points")
    if not isinstance(polygons, Polygons):
        raise ValueError("`polygons` must contain only polygons")
    if polygons.geometry_offset is not None and \
            polygons.geometry_offset.numel() != polygons.part_offset.numel():
        raise ValueError("GeoSeries cannot contain multipolygon.")
    dev = points.device
    res = _capi.TorchResource(dev)
    out = _vp()
    n = ctypes.c_int64()
    x = points.x.contiguous()
    y = points.y.contiguous()
    fn = (_capi.lib().qpip_pairwise_point_in_polygon if pairwise
          else _capi.lib().qpip_point_in_polygon)
    _capi.check(fn(_p(x), _p(y), x.numel(), y.numel(), _p(polygons.part_offset),
                   polygons.part_offset.numel(), _p(polygons.ring_offset),
                   polygons.ring_offset.numel(), _p(polygons.x), _p(polygons.y),
                   polygons.x.numel(), polygons.y.numel(), ctypes.byref(res.struct),
                   _capi.stream_handle(dev), ctypes.byref(out), ctypes.byref(n)))
    return res.take(out.value, n.value, torch.uint8 if pairwise else torch.int32)


def point_in_polygon(points: Points, polygons: Polygons):
    """Which points fall within which polygons, every point against every
    polygon (join.py:23-101); at most 31 polygons. Returns a Table with one bool
    column per polygon (keyed by polygon position), one row per point."""
    if len(polygons) == 0:
        return Table()
    bits = _brute_force(points, polygons, pairwise=False)
    out = Table()
    for j in range(len(polygons)):
        out[j] = ((bits >> j) & 1).to(torch.bool)
    return out


def point_in_polygon_bitmask(points: Points, polygons: Polygons) -> torch.Tensor:
    """The INT32 bitmask column of cuspatial::point_in_polygon (bit j: polygon j)."""
    return _brute_force(points, polygons, pairwise=False)


def pairwise_point_in_polygon(points: Points, polygons: Polygons) -> torch.Tensor:
    """Point i against polygon i (cuspatial::pairwise_point_in_polygon); bool per pair."""
    return _brute_force(points, polygons, pairwise=True).to(torch.bool)


# ---------------------------------------------------------------------------
# Batched jobs
# ---------------------------------------------------------------------------
class BatchResult:
    """Results of batched_quadtree_point_in_polygon. Job j's point_indices are
    point_indices[job_point_offsets[j]:job_point_offsets[j+1]] and its hits
    polygon_index / point_index[job_hit_offsets[j]:job_hit_offsets[j+1]], all with
    job-local values, exactly as the single-job calls return them."""

    def __init__(self, point_indices, polygon_index, point_index, job_point_offsets,
                 job_hit_offsets):
        self.point_indices = point_indices
        self.polygon_index = polygon_index
        self.point_index = point_index
        self.job_point_offsets = job_point_offsets
        self.job_hit_offsets = job_hit_offsets

    def __len__(self):
        return len(self.job_hit_offsets) - 1

    @property
    def num_hits(self):
        return int(self.job_hit_offsets[-1])

    def job(self, j):
        """(point_indices, Table(polygon_index, point_index)) of job j."""
        p0, p1 = int(self.job_point_offsets[j]), int(self.job_point_offsets[j + 1])
        h0, h1 = int(self.job_hit_offsets[j]), int(self.job_hit_offsets[j + 1])
        out = Table()
        out["polygon_index"] = self.polygon_index[h0:h1]
        out["point_index"] = self.point_index[h0:h1]
        return self.point_indices[p0:p1], out


def batched_quadtree_point_in_polygon(points: Points, job_point_offsets, polygons: Polygons,
                                      job_polygon_offsets, bboxes, scales, max_depths,
                                      max_sizes) -> BatchResult:
    """quadtree_on_points -> polygon_bounding_boxes -> join_quadtree_and_bounding_boxes
    -> quadtree_point_in_polygon for many independent jobs in one pass of each
    stage. Job j owns points [job_point_offsets[j], job_point_offsets[j+1]) and
    polygons [job_polygon_offsets[j], job_polygon_offsets[j+1]); bboxes holds
    (x_min, x_max, y_min, y_max) per job. Per job the result equals the four
    single-job calls with the same parameters (bbox ordering and scale clamp
    included; the clamp warning is not repeated per job)."""
    if not isinstance(points, Points):
        raise ValueError("GeoSeries must contain only points.")
    if not isinstance(polygons, Polygons):
        raise ValueError("`polygons` Geoseries must contains only polygons geometries.")
    jpo = np.ascontiguousarray(np.asarray(job_point_offsets, dtype=np.int64))
    jpy = np.ascontiguousarray(np.asarray(job_polygon_offsets, dtype=np.int64))
    J = len(jpo) - 1
    if len(jpy) != J + 1:
        raise ValueError("job_point_offsets and job_polygon_offsets must have the same length")
    bb = np.ascontiguousarray(np.asarray(bboxes, dtype=np.float64).reshape(J, 4))
    sc = np.ascontiguousarray(np.asarray(scales, dtype=np.float64).reshape(J))
    md = np.ascontiguousarray(np.asarray(max_depths, dtype=np.int8).reshape(J))
    ms = np.ascontiguousarray(np.asarray(max_sizes, dtype=np.int32).reshape(J))
    dev = points.device
    xy = points.xy if points.xy is not None else \
        torch.stack([points.x, points.y], dim=1).reshape(-1).contiguous()
    res = _capi.TorchResource(dev)
    pi, po, pt = _outs(3)
    hits = np.zeros(J + 1, dtype=np.int64)
    h = ctypes.c_int64()
    ptr = lambda a: a.ctypes.data_as(_vp)  # noqa: E731
    _capi.check(_capi.lib().qpip_batched_quadtree_point_in_polygon(
        J, ptr(jpo), ptr(jpy), ptr(bb), ptr(sc), ptr(md), ptr(ms), _p(xy), len(points),
        _p(polygons.part_offset), polygons.part_offset.numel(), _p(polygons.ring_offset),
        polygons.ring_offset.numel(), _p(polygons.x), _p(polygons.y), polygons.x.numel(),
        polygons.y.numel(), ctypes.byref(res.struct), _capi.stream_handle(dev), ctypes.byref(pi),
        ctypes.byref(po), ctypes.byref(pt), ptr(hits), ctypes.byref(h)))
    n = len(points)
    return BatchResult(res.take(pi.value, n if J > 0 else 0, torch.uint32),
                       res.take(po.value, h.value, torch.uint32),
                       res.take(pt.value, h.value, torch.uint32), jpo, hits)


class BatchPlan:
    """The batched join captured once as a HIP graph and replayed per launch.

    The plan fixes the jobs, the polygons and the point buffer (``points.xy``,
    interleaved on the device): ``launch()`` replays the whole pipeline over the
    points that buffer holds at that moment, with no host synchronisation and no
    allocation, and ``result()`` returns what ``batched_quadtree_point_in_polygon``
    returns for those points. Rewrite the buffer in place between launches to
    join new points with the same job layout. Capacities (leaf pairs, PIP chunks,
    hits) are sized from the points present at creation times ``headroom``; a
    launch that needs more raises ``CapacityError`` from ``result()`` and writes
    nothing out of bounds.
    """

    def __init__(self, points: Points, job_point_offsets, polygons: Polygons,
                 job_polygon_offsets, bboxes, scales, max_depths, max_sizes,
                 headroom: float = 1.25, tuning="default"):
        """`tuning`: "default" looks the batch's shape bucket up in the package's
        tuning table (autotune.DEFAULT_TABLE); a TuningTable or a path looks it up
        there; a dict is used as is; None captures under the current values. The
        values apply while the plan is sized and captured, and are restored
        after."""
        from . import autotune

        ring_n = polygons.ring_offset.numel() - 1
        edges = max(polygons.x.numel() - max(ring_n, 0), 0)
        self.bucket = autotune.shape_bucket(len(job_point_offsets) - 1, len(points),
                                            max(len(polygons), 1), edges)
        if isinstance(tuning, dict) or tuning is None:
            cfg = tuning
        else:
            table = tuning if isinstance(tuning, autotune.TuningTable) else autotune.TuningTable(
                autotune.DEFAULT_TABLE if tuning == "default" else tuning)
            cfg = table.lookup(self.bucket)
        self.tuning = dict(cfg) if cfg else None
        with autotune.applied(self.tuning):
            self._create(points, job_point_offsets, polygons, job_polygon_offsets, bboxes,
                         scales, max_depths, max_sizes, headroom)

    def _create(self, points, job_point_offsets, polygons, job_polygon_offsets, bboxes, scales,
                max_depths, max_sizes, headroom):
        if not isinstance(points, Points) or points.xy is None:
            raise ValueError("a plan needs interleaved points (Points.from_points_xy)")
        if not isinstance(polygons, Polygons):
            raise ValueError("`polygons` Geoseries must contains only polygons geometries.")
        self.jpo = np.ascontiguousarray(np.asarray(job_point_offsets, dtype=np.int64))
        jpy = np.ascontiguousarray(np.asarray(job_polygon_offsets, dtype=np.int64))
        J = len(self.jpo) - 1
        if len(jpy) != J + 1:
            raise ValueError("job_point_offsets and job_polygon_offsets must have the same length")
        bb = np.ascontiguousarray(np.asarray(bboxes, dtype=np.float64).reshape(J, 4))
        sc = np.ascontiguousarray(np.asarray(scales, dtype=np.float64).reshape(J))
        md = np.ascontiguousarray(np.asarray(max_depths, dtype=np.int8).reshape(J))
        ms = np.ascontiguousarray(np.asarray(max_sizes, dtype=np.int32).reshape(J))
        self.points, self.polygons = points, polygons  # kept alive: the graph reads them
        self.device = points.device
        self.num_jobs, self.num_points = J, len(points)
        h = ctypes.c_void_p()
        ptr = lambda a: a.ctypes.data_as(_vp)  # noqa: E731
        _capi.check(_capi.lib().qpip_batch_plan_create(
            J, ptr(self.jpo), ptr(jpy), ptr(bb), ptr(sc), ptr(md), ptr(ms), _p(points.xy),
            len(points), _p(polygons.part_offset), polygons.part_offset.numel(),
            _p(polygons.ring_offset), polygons.ring_offset.numel(), _p(polygons.x),
            _p(polygons.y), polygons.x.numel(), polygons.y.numel(), float(headroom),
            _capi.stream_handle(self.device), ctypes.byref(h)))
        self._h = h

    def launch(self) -> None:
        _capi.check(_capi.lib().qpip_batch_plan_launch(self._h,
                                                       _capi.stream_handle(self.device)))

    def result(self) -> BatchResult:
        """Waits for the last launch; its outputs, copied into torch tensors."""
        res = _capi.TorchResource(self.device)
        pi, po, pt = _outs(3)
        hits = np.zeros(self.num_jobs + 1, dtype=np.int64)
        h = ctypes.c_int64()
        _capi.check(_capi.lib().qpip_batch_plan_result(
            self._h, ctypes.byref(res.struct), _capi.stream_handle(self.device),
            ctypes.byref(pi), ctypes.byref(po), ctypes.byref(pt), hits.ctypes.data_as(_vp),
            ctypes.byref(h)))
        return BatchResult(res.take(pi.value, self.num_points, torch.uint32),
                           res.take(po.value, h.value, torch.uint32),
                           res.take(pt.value, h.value, torch.uint32), self.jpo, hits)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _capi.check(_capi.lib().qpip_batch_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


# ---------------------------------------------------------------------------
# Linestrings
# ---------------------------------------------------------------------------
def linestring_bounding_boxes(linestrings: Linestrings, expansion_radius: float):
    """Bounding box of each linestring, every vertex expanded by expansion_radius
    (bounding.py:83-135). For multilinestrings the box covers all parts."""
    cols = ["minx", "miny", "maxx", "maxy"]
    if len(linestrings) == 0:
        out = Table()
        for c in cols:
            out[c] = torch.zeros(0, dtype=torch.float64, device=linestrings.device)
        return out
    if not isinstance(linestrings, Linestrings):
        raise ValueError("Geoseries must contain only linestrings.")
    offs = linestrings.part_offset
    if linestrings.geometry_offset is not None:  # part_offset.take(geometry_offset)
        offs = offs[linestrings.geometry_offset.to(torch.int64)].contiguous()
    dev = linestrings.device
    res = _capi.TorchResource(dev)
    o = _outs(4)
    n = ctypes.c_int64()
    _capi.check(_capi.lib().qpip_linestring_bounding_boxes(
        _p(offs), offs.numel(), _p(linestrings.x), _p(linestrings.y), linestrings.x.numel(),
        linestrings.y.numel(), float(expansion_radius), ctypes.byref(res.struct),
        _capi.stream_handle(dev), *[ctypes.byref(v) for v in o], ctypes.byref(n)))
    out = Table()
    for c, v in zip(cols, o):
        out[c] = res.take(v.value, n.value, torch.float64)
    return out


def quadtree_point_to_nearest_linestring(linestring_quad_pairs, quadtree, point_indices,
                                         points: Points, linestrings: Linestrings):
    """Nearest linestring of each point among those paired with its quadrant
    (join.py:265-355). Returns columns point_index (a position in
    ``point_indices``), linestring_index and distance, one row per point."""
    if not isinstance(points, Points):
        raise ValueError("`point` Geoseries must contains only point geometries.")
    if not isinstance(linestrings, Linestrings):
        raise ValueError("`linestrings` Geoseries must contains only linestring geometries.")
    if linestrings.geometry_offset is not None and \
            linestrings.geometry_offset.numel() != linestrings.part_offset.numel():
        raise ValueError("GeoSeries cannot contain multilinestrings.")
    pcols = list(linestring_quad_pairs.values()) if isinstance(linestring_quad_pairs, dict) \
        else list(linestring_quad_pairs)
    if len(pcols) != 2:
        raise RuntimeError("cuSpatial failure: a quadrant-linestring table must have 2 columns")
    key, level, internal, length, offset = _tree_columns(quadtree)
    dev = points.device
    pl = _u32(pcols[0], "linestring_quad_pairs")
    pq = _u32(pcols[1], "linestring_quad_pairs")
    pi = _u32(point_indices, "point_indices")
    x = points.x.contiguous()
    y = points.y.contiguous()
    lo = _u32(linestrings.part_offset, "linestring_offsets")
    res = _capi.TorchResource(dev)
    o = _outs(3)
    n = ctypes.c_int64()
    _capi.check(_capi.lib().qpip_quadtree_point_to_nearest_linestring(
        _p(pl), _p(pq), pl.numel(), _p(key), _p(level), _p(internal), _p(length), _p(offset),
        key.numel(), _p(pi), pi.numel(), _p(x), _p(y), x.numel(), y.numel(), _p(lo), lo.numel(),
        _p(linestrings.x), _p(linestrings.y), linestrings.x.numel(), linestrings.y.numel(),
        ctypes.byref(res.struct), _capi.stream_handle(dev), *[ctypes.byref(v) for v in o],
        ctypes.byref(n)))
    out = Table()
    out["point_index"] = res.take(o[0].value, n.value, torch.uint32)
    out["linestring_index"] = res.take(o[1].value, n.value, torch.uint32)
    out["distance"] = res.take(o[2].value, n.value, torch.float64)
    return out
