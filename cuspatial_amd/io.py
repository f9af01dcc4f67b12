"""GeoArrow input/output for points and polygons, and a versioned file format for
a resident index (tree, polygons, boxes and any extra arrays, such as ranker
weights), for restarts of a long-running service.

Stands in for the reference's host-side readers: ``from_geopandas``
(python/cuspatial/cuspatial/io/geopandas_reader.py) and its arrow layouts
(python/cuspatial/cuspatial/io/pygeoarrow.py:7-16). Geometries are read from, and
written as, GeoArrow's native encodings, which carry the same offsets the
quadtree path consumes:

  point         fixed_size_list<double>[2]  (interleaved x, y; "geoarrow.point")
                or struct<x: double, y: double> (separated coordinates)
  polygon       list<list<coord>>            (rings of a polygon)
  multipolygon  list<list<list<coord>>>      (parts, rings)

Reading also accepts the reference's own layouts: coordinates as list<double>
of two values, and the dense union of pygeoarrow.py:19-52 (its "points" and
"polygons" children). Arrays move to the device with one copy each; the offset
buffers are used as stored (rebased when a slice starts past zero).
"""

from __future__ import annotations

import json

import numpy as np
import pyarrow as pa
import pyarrow.ipc as ipc
import pyarrow.parquet as pq
import torch

from .spatial import Points, Polygons

COORD = pa.list_(pa.float64(), 2)
POINT_EXT = "geoarrow.point"
POLYGON_EXT = "geoarrow.polygon"
MULTIPOLYGON_EXT = "geoarrow.multipolygon"
INDEX_FORMAT = "cuspatial_amd.index"
INDEX_VERSION = 1


# ---------------------------------------------------------------------------
# Arrow -> Points / Polygons
# ---------------------------------------------------------------------------
def _combine(arr):
    if isinstance(arr, pa.ChunkedArray):
        arr = arr.combine_chunks() if arr.num_chunks != 1 else arr.chunk(0)
    if isinstance(arr, pa.ExtensionArray):
        arr = arr.storage
    return arr


def _offsets(list_arr) -> np.ndarray:
    """Offsets of a list array, rebased to start at 0 (a slice keeps the parent's)."""
    off = list_arr.offsets.to_numpy(zero_copy_only=False).astype(np.int64)
    return off - off[0] if off.size else np.zeros(1, np.int64)


def _flat_values(list_arr):
    """The child values a list array's offsets point into, sliced to this array."""
    off = list_arr.offsets.to_numpy(zero_copy_only=False)
    if off.size == 0:
        return list_arr.values.slice(0, 0)
    return list_arr.values.slice(int(off[0]), int(off[-1] - off[0]))


def _coords(arr) -> np.ndarray:
    """Coordinates (n, 2) float64 from a coordinate array in any accepted encoding."""
    t = arr.type
    if pa.types.is_fixed_size_list(t):
        if t.list_size != 2:
            raise ValueError("only 2-dimensional (x, y) coordinates are supported")
        vals = arr.values.slice(arr.offset * 2, len(arr) * 2)
        return vals.to_numpy(zero_copy_only=False).astype(np.float64).reshape(-1, 2)
    if pa.types.is_struct(t):
        names = [t.field(i).name for i in range(t.num_fields)]
        if names[:2] != ["x", "y"]:
            raise ValueError("struct coordinates must start with fields x, y")
        x = arr.field(0).to_numpy(zero_copy_only=False).astype(np.float64)
        y = arr.field(1).to_numpy(zero_copy_only=False).astype(np.float64)
        return np.stack([x, y], 1)
    if pa.types.is_list(t) or pa.types.is_large_list(t):  # pygeoarrow: list<double> of 2
        off = _offsets(arr)
        if off.size > 1 and not np.all(np.diff(off) == 2):
            raise ValueError("coordinates must hold exactly two values")
        return _flat_values(arr).to_numpy(zero_copy_only=False).astype(np.float64).reshape(-1, 2)
    raise TypeError(f"not a coordinate array: {t}")


def _union_child(arr, name):
    if not pa.types.is_union(arr.type):
        return arr
    names = [arr.type.field(i).name for i in range(arr.type.num_fields)]
    if name not in names:
        raise TypeError(f"union has no '{name}' child")
    k = names.index(name)
    codes = arr.type.type_codes
    types = arr.type_codes.to_numpy(zero_copy_only=False)
    if np.any(types != codes[k]):
        raise TypeError(f"union holds geometries other than {name}")
    child = arr.field(k)
    offs = arr.offsets.to_numpy(zero_copy_only=False)
    if offs.size and not np.array_equal(offs, np.arange(offs[0], offs[0] + offs.size)):
        child = child.take(pa.array(offs))
    elif offs.size:
        child = child.slice(int(offs[0]), offs.size)
    else:
        child = child.slice(0, 0)
    return child


def points_from_arrow(arr, device=None) -> Points:
    """Points from a GeoArrow point array (interleaved or separated), a
    reference-layout list<double> point array, or a union holding only points."""
    arr = _union_child(_combine(arr), "points")
    if arr.null_count:
        raise ValueError("null points are not supported")
    xy = _coords(arr)
    t = torch.from_numpy(np.ascontiguousarray(xy).reshape(-1))
    return Points.from_points_xy(t if device is None else t.to(device), device=device)


def polygons_from_arrow(arr, device=None) -> Polygons:
    """Polygons from a GeoArrow polygon or multipolygon array, the reference's
    multipolygon-shaped list layout, or a union holding only polygons. A
    multipolygon column keeps its geometry offsets (geometry_offset)."""
    arr = _union_child(_combine(arr), "polygons")
    if arr.null_count:
        raise ValueError("null polygons are not supported")
    depth, t = 0, arr.type
    while pa.types.is_list(t) or pa.types.is_large_list(t):
        depth += 1
        t = t.value_type
    if pa.types.is_fixed_size_list(t) or pa.types.is_struct(t):
        coord_lists = depth  # list levels above a coordinate
    elif pa.types.is_floating(t):
        coord_lists = depth - 1  # innermost list<double> is the coordinate
    else:
        raise TypeError(f"not a polygon array: {arr.type}")
    geometry_offset = None
    if coord_lists == 3:  # multipolygon: geometries -> parts -> rings -> coords
        geometry_offset = _offsets(arr)
        arr = _flat_values(arr)
    elif coord_lists != 2:
        raise TypeError(f"not a polygon array: {arr.type}")
    part = _offsets(arr)
    rings = _flat_values(arr)
    ring = _offsets(rings)
    xy = _coords(_flat_values(rings))
    return Polygons(part.astype(np.int32), ring.astype(np.int32),
                    np.ascontiguousarray(xy[:, 0]), np.ascontiguousarray(xy[:, 1]),
                    geometry_offset=None if geometry_offset is None else
                    geometry_offset.astype(np.int32), device=device)


# ---------------------------------------------------------------------------
# Points / Polygons -> Arrow
# ---------------------------------------------------------------------------
def _np(t) -> np.ndarray:
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def points_to_arrow(points: Points) -> pa.Array:
    """GeoArrow point array with interleaved coordinates."""
    if points.xy is not None:
        flat = _np(points.xy).astype(np.float64).reshape(-1)
    else:
        flat = np.stack([_np(points.x), _np(points.y)], 1).astype(np.float64).reshape(-1)
    return pa.FixedSizeListArray.from_arrays(pa.array(flat, pa.float64()), 2)


def polygons_to_arrow(polygons: Polygons) -> pa.Array:
    """GeoArrow polygon array, or multipolygon when geometry_offset is set."""
    xy = np.stack([_np(polygons.x), _np(polygons.y)], 1).astype(np.float64).reshape(-1)
    coords = pa.FixedSizeListArray.from_arrays(pa.array(xy, pa.float64()), 2)
    rings = pa.ListArray.from_arrays(pa.array(_np(polygons.ring_offset).astype(np.int32)), coords)
    arr = pa.ListArray.from_arrays(pa.array(_np(polygons.part_offset).astype(np.int32)), rings)
    if polygons.geometry_offset is not None:
        arr = pa.ListArray.from_arrays(
            pa.array(_np(polygons.geometry_offset).astype(np.int32)), arr)
    return arr


def _field(name, arr, ext):
    return pa.field(name, arr.type, metadata={b"ARROW:extension:name": ext.encode(),
                                              b"ARROW:extension:metadata": b"{}"})


def to_table(**columns) -> pa.Table:
    """A table of geometry columns (Points or Polygons), tagged with GeoArrow
    extension names in the field metadata."""
    arrays, fields = [], []
    for name, g in columns.items():
        if isinstance(g, Points):
            a, ext = points_to_arrow(g), POINT_EXT
        elif isinstance(g, Polygons):
            a = polygons_to_arrow(g)
            ext = MULTIPOLYGON_EXT if g.geometry_offset is not None else POLYGON_EXT
        else:
            raise TypeError(f"column {name}: expected Points or Polygons")
        arrays.append(a)
        fields.append(_field(name, a, ext))
    return pa.Table.from_arrays(arrays, schema=pa.schema(fields))


def from_table(table: pa.Table, device=None) -> dict:
    """Points / Polygons per geometry column of a table (by extension name when
    present, else by shape)."""
    out = {}
    for i, f in enumerate(table.schema):
        ext = (f.metadata or {}).get(b"ARROW:extension:name", b"").decode()
        col = table.column(i)
        if ext == POINT_EXT:
            out[f.name] = points_from_arrow(col, device)
        elif ext in (POLYGON_EXT, MULTIPOLYGON_EXT):
            out[f.name] = polygons_from_arrow(col, device)
        else:
            try:
                out[f.name] = points_from_arrow(col, device)
            except (TypeError, ValueError):
                out[f.name] = polygons_from_arrow(col, device)
    return out


def write_ipc(path, **columns) -> None:
    t = to_table(**columns)
    with pa.OSFile(str(path), "wb") as sink, ipc.new_file(sink, t.schema) as w:
        w.write_table(t)


def read_ipc(path, device=None) -> dict:
    with pa.memory_map(str(path), "r") as src:
        return from_table(ipc.open_file(src).read_all(), device)


def write_parquet(path, **columns) -> None:
    pq.write_table(to_table(**columns), str(path))


def read_parquet(path, device=None) -> dict:
    return from_table(pq.read_table(str(path)), device)


# ---------------------------------------------------------------------------
# Versioned index file
# ---------------------------------------------------------------------------
def save_index(path, arrays: dict, meta: dict | None = None) -> None:
    """Write named arrays (numpy or torch, any length) and JSON metadata to an
    Arrow IPC file: one row, one large_list column per array, the element type
    kept; the schema metadata holds the format name, version and ``meta``."""
    cols, fields = [], []
    for name, a in arrays.items():
        v = _np(a)
        if v.dtype == np.bool_:
            v = v.astype(np.uint8)
        flat = pa.array(np.ascontiguousarray(v).reshape(-1))
        col = pa.LargeListArray.from_arrays(pa.array([0, len(flat)], pa.int64()), flat)
        cols.append(col)
        fields.append(pa.field(name, col.type, metadata={
            b"shape": json.dumps(list(v.shape)).encode(), b"dtype": str(_np(a).dtype).encode()}))
    md = {b"format": INDEX_FORMAT.encode(), b"version": str(INDEX_VERSION).encode(),
          b"meta": json.dumps(meta or {}).encode()}
    t = pa.Table.from_arrays(cols, schema=pa.schema(fields, metadata=md))
    with pa.OSFile(str(path), "wb") as sink, ipc.new_file(sink, t.schema) as w:
        w.write_table(t)


def load_index(path, device=None):
    """(arrays, meta) from save_index. Arrays come back as torch tensors on
    ``device`` (numpy arrays when device is None). Raises ValueError for another
    format or a newer version."""
    with pa.memory_map(str(path), "r") as src:
        t = ipc.open_file(src).read_all()
    md = t.schema.metadata or {}
    if md.get(b"format", b"").decode() != INDEX_FORMAT:
        raise ValueError(f"{path}: not a {INDEX_FORMAT} file")
    version = int(md.get(b"version", b"0"))
    if version > INDEX_VERSION or version < 1:
        raise ValueError(f"{path}: index format version {version} is not supported "
                         f"(this build reads 1..{INDEX_VERSION})")
    arrays = {}
    for i, f in enumerate(t.schema):
        col = _combine(t.column(i))
        shape = json.loads(f.metadata[b"shape"])
        dtype = np.dtype(f.metadata[b"dtype"].decode())
        v = _flat_values(col).to_numpy(zero_copy_only=False)
        v = v.astype(dtype, copy=False).reshape(shape)
        if device is not None:
            v = torch.from_numpy(np.array(v, copy=True)).to(device)  # memory-mapped: read-only
        arrays[f.name] = v
    return arrays, json.loads(md.get(b"meta", b"{}"))
