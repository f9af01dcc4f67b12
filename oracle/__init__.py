"""ctypes loader for the oracle (test infrastructure only).

The oracle is a host C++ restatement of the reference quadtree point-in-polygon
path (``oracle/qpip_oracle.cpp``). Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` import this module, and only as the
checker or the reported CPU baseline; the product package never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libqpip_oracle.so")
_lib = None

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_u32_p = ctypes.POINTER(ctypes.c_uint32)
_c_i32_p = ctypes.POINTER(ctypes.c_int32)
_c_u8_p = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> str:
    """Compile the oracle with ``make`` (a few seconds)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_native(out_dir: str) -> str:
    """Compile the oracle for this host's CPU (``-O3 -march=native``) into
    out_dir and load that copy instead of the portable build (the CPU baseline
    runs on the GPU machine's host, whose CPU is only known there). The
    arithmetic is the same: floating-point contraction stays off."""
    global _lib, _LIB_PATH
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libqpip_oracle_native.so")
    cmd = [os.environ.get("CXX", "g++"), "-O3", "-march=native", "-mtune=native", "-fopenmp",
           "-ffp-contract=off", "-fPIC", "-std=c++17", "-shared",
           os.path.join(_HERE, "qpip_oracle.cpp"), "-o", out + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    _LIB_PATH = out
    _lib = None
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        i64 = ctypes.c_int64
        d = ctypes.c_double
        L.oq_result_status.argtypes = [vp]
        L.oq_result_status.restype = ctypes.c_int
        L.oq_result_error.argtypes = [vp]
        L.oq_result_error.restype = ctypes.c_char_p
        L.oq_result_num_arrays.argtypes = [vp]
        L.oq_result_num_arrays.restype = ctypes.c_int
        L.oq_result_count.argtypes = [vp, ctypes.c_int]
        L.oq_result_count.restype = i64
        L.oq_result_data.argtypes = [vp, ctypes.c_int]
        L.oq_result_data.restype = vp
        L.oq_result_free.argtypes = [vp]
        L.oq_set_threads.argtypes = [ctypes.c_int]
        L.oq_max_threads.restype = ctypes.c_int
        L.oq_z_order.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.oq_z_order.restype = ctypes.c_uint32
        L.oq_z_order_x.argtypes = [ctypes.c_uint32]
        L.oq_z_order_x.restype = ctypes.c_uint16
        L.oq_z_order_y.argtypes = [ctypes.c_uint32]
        L.oq_z_order_y.restype = ctypes.c_uint16
        L.oq_float_equal.argtypes = [d, d]
        L.oq_float_equal.restype = ctypes.c_int
        L.oq_is_point_in_polygon.argtypes = [d, d, _c_u32_p, i64, i64, _c_double_p, _c_double_p]
        L.oq_is_point_in_polygon.restype = ctypes.c_int
        L.oq_point_in_polygon_bitmask.argtypes = [
            _c_double_p, _c_double_p, i64, _c_u32_p, i64, _c_u32_p, _c_double_p, _c_double_p, _c_i32_p]
        L.oq_pairwise_point_in_polygon.argtypes = [
            _c_double_p, _c_double_p, i64, _c_u32_p, _c_u32_p, _c_double_p, _c_double_p, _c_u8_p]
        L.oq_all_pairs_bruteforce.argtypes = [
            _c_double_p, _c_double_p, i64, _c_u32_p, i64, _c_u32_p, _c_double_p, _c_double_p]
        L.oq_all_pairs_bruteforce.restype = vp
        L.oq_quadtree_on_points.argtypes = [
            _c_double_p, _c_double_p, i64, i64, d, d, d, d, d, ctypes.c_int32, ctypes.c_int32]
        L.oq_quadtree_on_points.restype = vp
        L.oq_polygon_bounding_boxes.argtypes = [
            _c_i32_p, i64, _c_i32_p, i64, _c_double_p, _c_double_p, i64, i64, d]
        L.oq_polygon_bounding_boxes.restype = vp
        L.oq_join_quadtree_and_bounding_boxes.argtypes = [
            _c_u32_p, _c_u8_p, _c_u8_p, _c_u32_p, _c_u32_p, i64,
            _c_double_p, _c_double_p, _c_double_p, _c_double_p, i64,
            d, d, d, d, d, ctypes.c_int32]
        L.oq_join_quadtree_and_bounding_boxes.restype = vp
        L.oq_quadtree_point_in_polygon.argtypes = [
            _c_u32_p, _c_u32_p, i64, _c_u32_p, _c_u32_p, i64, _c_u32_p, i64,
            _c_double_p, _c_double_p, i64, i64, _c_u32_p, i64, _c_u32_p, i64,
            _c_double_p, _c_double_p, i64, i64]
        L.oq_quadtree_point_in_polygon.restype = vp
        L.oq_quadtree_point_to_nearest_linestring.argtypes = [
            _c_u32_p, _c_u32_p, i64, _c_u32_p, _c_u32_p, i64, _c_u32_p, i64,
            _c_double_p, _c_double_p, i64, i64, _c_u32_p, i64,
            _c_double_p, _c_double_p, i64, i64]
        L.oq_quadtree_point_to_nearest_linestring.restype = vp
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _take(h, dtypes):
    L = lib()
    try:
        if L.oq_result_status(h) != 0:
            raise OracleError(L.oq_result_error(h).decode())
        out = []
        for i, dt in enumerate(dtypes):
            n = L.oq_result_count(h, i)
            if n == 0:
                out.append(np.zeros(0, dtype=dt))
                continue
            buf = (ctypes.c_char * (n * np.dtype(dt).itemsize)).from_address(L.oq_result_data(h, i))
            out.append(np.frombuffer(buf, dtype=dt).copy())
        return out
    finally:
        L.oq_result_free(h)


def set_threads(n: int) -> None:
    lib().oq_set_threads(int(n))


def max_threads() -> int:
    return int(lib().oq_max_threads())


def z_order(x: int, y: int) -> int:
    return int(lib().oq_z_order(x, y))


def z_order_x(k: int) -> int:
    return int(lib().oq_z_order_x(k))


def z_order_y(k: int) -> int:
    return int(lib().oq_z_order_y(k))


def float_equal(a: float, b: float) -> bool:
    return bool(lib().oq_float_equal(a, b))


def is_point_in_polygon(px, py, ring_offsets, ring_begin, ring_end, vx, vy) -> bool:
    ro = _arr(ring_offsets, np.uint32)
    vx = _arr(vx, np.float64)
    vy = _arr(vy, np.float64)
    return bool(lib().oq_is_point_in_polygon(
        float(px), float(py), _ptr(ro, ctypes.c_uint32), int(ring_begin), int(ring_end),
        _ptr(vx, ctypes.c_double), _ptr(vy, ctypes.c_double)))


def point_in_polygon_bitmask(x, y, poly_offsets, ring_offsets, vx, vy):
    x = _arr(x, np.float64)
    y = _arr(y, np.float64)
    po = _arr(poly_offsets, np.uint32)
    ro = _arr(ring_offsets, np.uint32)
    vx = _arr(vx, np.float64)
    vy = _arr(vy, np.float64)
    out = np.zeros(len(x), dtype=np.int32)
    lib().oq_point_in_polygon_bitmask(
        _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), len(x), _ptr(po, ctypes.c_uint32),
        len(po), _ptr(ro, ctypes.c_uint32), _ptr(vx, ctypes.c_double), _ptr(vy, ctypes.c_double),
        _ptr(out, ctypes.c_int32))
    return out


def pairwise_point_in_polygon(x, y, poly_offsets, ring_offsets, vx, vy):
    """Point i against polygon i (0/1 per pair)."""
    x = _arr(x, np.float64)
    y = _arr(y, np.float64)
    po = _arr(poly_offsets, np.uint32)
    ro = _arr(ring_offsets, np.uint32)
    vx = _arr(vx, np.float64)
    vy = _arr(vy, np.float64)
    assert len(po) == len(x) + 1
    out = np.zeros(len(x), dtype=np.uint8)
    lib().oq_pairwise_point_in_polygon(
        _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), len(x), _ptr(po, ctypes.c_uint32),
        _ptr(ro, ctypes.c_uint32), _ptr(vx, ctypes.c_double), _ptr(vy, ctypes.c_double),
        _ptr(out, ctypes.c_uint8))
    return out


def all_pairs_bruteforce(x, y, poly_offsets, ring_offsets, vx, vy):
    """(polygon_index, original point index) of every point strictly inside a polygon."""
    x = _arr(x, np.float64)
    y = _arr(y, np.float64)
    po = _arr(poly_offsets, np.uint32)
    ro = _arr(ring_offsets, np.uint32)
    vx = _arr(vx, np.float64)
    vy = _arr(vy, np.float64)
    h = lib().oq_all_pairs_bruteforce(
        _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), len(x), _ptr(po, ctypes.c_uint32),
        len(po), _ptr(ro, ctypes.c_uint32), _ptr(vx, ctypes.c_double), _ptr(vy, ctypes.c_double))
    return _take(h, [np.uint32, np.uint32])


def quadtree_on_points(x, y, x_min, x_max, y_min, y_max, scale, max_depth, max_size):
    """Returns (point_indices, key, level, is_internal_node, length, offset)."""
    x = _arr(x, np.float64)
    y = _arr(y, np.float64)
    h = lib().oq_quadtree_on_points(
        _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), len(x), len(y), float(x_min),
        float(x_max), float(y_min), float(y_max), float(scale), int(max_depth), int(max_size))
    pi, key, level, internal, length, offset = _take(
        h, [np.uint32, np.uint32, np.uint8, np.uint8, np.uint32, np.uint32])
    return pi, key, level, internal.astype(bool), length, offset


def polygon_bounding_boxes(poly_offsets, ring_offsets, x, y, expansion_radius=0.0):
    """Returns (minx, miny, maxx, maxy)."""
    po = _arr(poly_offsets, np.int32)
    ro = _arr(ring_offsets, np.int32)
    x = _arr(x, np.float64)
    y = _arr(y, np.float64)
    h = lib().oq_polygon_bounding_boxes(
        _ptr(po, ctypes.c_int32), len(po), _ptr(ro, ctypes.c_int32), len(ro),
        _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), len(x), len(y),
        float(expansion_radius))
    return tuple(_take(h, [np.float64] * 4))


def join_quadtree_and_bounding_boxes(tree, bboxes, x_min, x_max, y_min, y_max, scale, max_depth):
    """tree = (key, level, is_internal_node, length, offset); bboxes = (minx, miny, maxx, maxy).

    Returns (bbox_offset, quad_offset)."""
    key, level, internal, length, offset = tree
    key = _arr(key, np.uint32)
    level = _arr(level, np.uint8)
    internal = _arr(internal, np.uint8)
    length = _arr(length, np.uint32)
    offset = _arr(offset, np.uint32)
    bx = [_arr(b, np.float64) for b in bboxes]
    h = lib().oq_join_quadtree_and_bounding_boxes(
        _ptr(key, ctypes.c_uint32), _ptr(level, ctypes.c_uint8), _ptr(internal, ctypes.c_uint8),
        _ptr(length, ctypes.c_uint32), _ptr(offset, ctypes.c_uint32), len(key),
        *[_ptr(b, ctypes.c_double) for b in bx], len(bx[0]),
        float(x_min), float(x_max), float(y_min), float(y_max), float(scale), int(max_depth))
    return tuple(_take(h, [np.uint32, np.uint32]))


def quadtree_point_in_polygon(pairs, tree, point_indices, x, y, poly_offsets, ring_offsets, vx, vy):
    """pairs = (bbox_offset, quad_offset). Returns (polygon_index, point_index)."""
    pp = _arr(pairs[0], np.uint32)
    pq = _arr(pairs[1], np.uint32)
    length = _arr(tree[3], np.uint32)
    offset = _arr(tree[4], np.uint32)
    pi = _arr(point_indices, np.uint32)
    x = _arr(x, np.float64)
    y = _arr(y, np.float64)
    po = _arr(poly_offsets, np.uint32)
    ro = _arr(ring_offsets, np.uint32)
    vx = _arr(vx, np.float64)
    vy = _arr(vy, np.float64)
    h = lib().oq_quadtree_point_in_polygon(
        _ptr(pp, ctypes.c_uint32), _ptr(pq, ctypes.c_uint32), len(pp),
        _ptr(length, ctypes.c_uint32), _ptr(offset, ctypes.c_uint32), len(length),
        _ptr(pi, ctypes.c_uint32), len(pi), _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double),
        len(x), len(y), _ptr(po, ctypes.c_uint32), len(po), _ptr(ro, ctypes.c_uint32), len(ro),
        _ptr(vx, ctypes.c_double), _ptr(vy, ctypes.c_double), len(vx), len(vy))
    return tuple(_take(h, [np.uint32, np.uint32]))


def linestring_bounding_boxes(linestring_offsets, x, y, expansion_radius=0.0):
    """Box per linestring (detail/bounding_boxes.cuh:97-131): the same per-geometry
    reduction as polygon_bounding_boxes with one ring per polygon. Returns
    (minx, miny, maxx, maxy)."""
    lo = _arr(linestring_offsets, np.int32)
    n = max(len(lo) - 1, 0)
    return polygon_bounding_boxes(np.arange(n + 1, dtype=np.int32), lo, x, y, expansion_radius)


def quadtree_point_to_nearest_linestring(pairs, tree, point_indices, x, y, linestring_offsets,
                                         lx, ly):
    """pairs = (linestring_index, quad_index). Returns (point_index, linestring_index,
    distance), one row per entry of point_indices."""
    pl = _arr(pairs[0], np.uint32)
    pq = _arr(pairs[1], np.uint32)
    length = _arr(tree[3], np.uint32)
    offset = _arr(tree[4], np.uint32)
    pi = _arr(point_indices, np.uint32)
    x = _arr(x, np.float64)
    y = _arr(y, np.float64)
    lo = _arr(linestring_offsets, np.uint32)
    lx = _arr(lx, np.float64)
    ly = _arr(ly, np.float64)
    h = lib().oq_quadtree_point_to_nearest_linestring(
        _ptr(pl, ctypes.c_uint32), _ptr(pq, ctypes.c_uint32), len(pl),
        _ptr(length, ctypes.c_uint32), _ptr(offset, ctypes.c_uint32), len(length),
        _ptr(pi, ctypes.c_uint32), len(pi), _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double),
        len(x), len(y), _ptr(lo, ctypes.c_uint32), len(lo), _ptr(lx, ctypes.c_double),
        _ptr(ly, ctypes.c_double), len(lx), len(ly))
    return tuple(_take(h, [np.uint32, np.uint32, np.float64]))
