"""ctypes binding of libcuspatial_amd (the C ABI declared in include/cuspatial_amd.h).

This is the counterpart of the reference's Cython layer
(python/cuspatial/cuspatial/_lib/quadtree.pyx, spatial_join.pyx,
polygon_bounding_boxes.pyx): it releases the GIL around each call (ctypes does)
and maps the library's status codes to the exceptions Cython's ``except +``
raises for the reference (logic_error and cuda_error -> RuntimeError,
out_of_memory -> MemoryError).

Outputs are allocated through a ``qpip_resource`` whose callbacks create torch
tensors, so every output column is a torch tensor on the input's device with
no copy.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuspatial_amd.so")

QPIP_SUCCESS = 0
QPIP_ERROR_LOGIC = 1
QPIP_ERROR_DEVICE = 2
QPIP_ERROR_OOM = 3
QPIP_ERROR_CAPACITY = 4


class CapacityError(RuntimeError):
    """A launch of a captured plan outgrew a capacity fixed when the plan was made."""

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double

_ALLOC = ctypes.CFUNCTYPE(_vp, _vp, ctypes.c_size_t, _vp)
_FREE = ctypes.CFUNCTYPE(None, _vp, _vp, ctypes.c_size_t, _vp)


class qpip_resource(ctypes.Structure):
    _fields_ = [("allocate", _ALLOC), ("deallocate", _FREE), ("ctx", _vp)]


_lib = None
_lib_lock = threading.Lock()

# Every symbol the C header declares (checked by the CPU test suite).
EXPORTED = (
    "qpip_last_error_message",
    "qpip_version",
    "qpip_free",
    "qpip_set_memory_limit",
    "qpip_trim",
    "qpip_memory_info",
    "qpip_profile_enable",
    "qpip_profile_reset",
    "qpip_profile_report",
    "qpip_tuning_set",
    "qpip_tuning_get",
    "qpip_quadtree_on_points",
    "qpip_quadtree_on_points_xy",
    "qpip_polygon_bounding_boxes",
    "qpip_join_quadtree_and_bounding_boxes",
    "qpip_quadtree_point_in_polygon",
    "qpip_quadtree_point_in_polygon_xy",
    "qpip_point_in_polygon",
    "qpip_point_in_polygon_noalloc",
    "qpip_pairwise_point_in_polygon",
    "qpip_batched_quadtree_point_in_polygon",
    "qpip_batch_plan_create",
    "qpip_batch_plan_launch",
    "qpip_batch_plan_result",
    "qpip_batch_plan_destroy",
    "qpip_linestring_bounding_boxes",
    "qpip_quadtree_point_to_nearest_linestring",
    "qpip_radix_sort_pairs",
    "qpip_compact_nonnegative",
    "qpip_ranker_num_params",
    "qpip_ranker_features",
    "qpip_ranker_logits",
    "qpip_ranker_gradients",
    "qpip_ranker_adam",
    "qpip_ranker_locate",
)


def lib():
    """Loads the in-tree library. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cuspatial_amd._build` "
                "(the HIP library is required; there is no fallback)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.qpip_last_error_message.restype = ctypes.c_char_p
        L.qpip_version.restype = ctypes.c_char_p
        L.qpip_free.argtypes = [_vp]
        L.qpip_set_memory_limit.argtypes = [ctypes.c_size_t]
        L.qpip_memory_info.argtypes = [P(ctypes.c_size_t), P(ctypes.c_size_t)]
        L.qpip_memory_info.restype = None
        L.qpip_profile_enable.argtypes = [ctypes.c_int]
        L.qpip_profile_report.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.qpip_profile_report.restype = ctypes.c_int64
        L.qpip_tuning_set.argtypes = [ctypes.c_char_p, _i64]
        L.qpip_tuning_set.restype = ctypes.c_int
        L.qpip_tuning_get.argtypes = [ctypes.c_char_p, P(_i64)]
        L.qpip_tuning_get.restype = ctypes.c_int
        L.qpip_quadtree_on_points.argtypes = [
            _vp, _vp, _i64, _i64, _d, _d, _d, _d, _d, ctypes.c_int8, ctypes.c_int32,
            P(qpip_resource), _vp, P(_vp), P(_vp), P(_vp), P(_vp), P(_vp), P(_vp), P(_i64)]
        L.qpip_quadtree_on_points_xy.argtypes = [
            _vp, _i64, _d, _d, _d, _d, _d, ctypes.c_int8, ctypes.c_int32,
            P(qpip_resource), _vp, P(_vp), P(_vp), P(_vp), P(_vp), P(_vp), P(_vp), P(_i64)]
        L.qpip_polygon_bounding_boxes.argtypes = [
            _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64, _d, P(qpip_resource), _vp,
            P(_vp), P(_vp), P(_vp), P(_vp), P(_i64)]
        L.qpip_join_quadtree_and_bounding_boxes.argtypes = [
            _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _d, _d, _d, _d, _d,
            ctypes.c_int8, P(qpip_resource), _vp, P(_vp), P(_vp), P(_i64)]
        L.qpip_quadtree_point_in_polygon.argtypes = [
            _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64,
            _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64, P(qpip_resource), _vp,
            P(_vp), P(_vp), P(_i64)]
        L.qpip_quadtree_point_in_polygon_xy.argtypes = [
            _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64,
            _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64, P(qpip_resource), _vp,
            P(_vp), P(_vp), P(_i64)]
        for name in ("qpip_point_in_polygon", "qpip_pairwise_point_in_polygon"):
            getattr(L, name).argtypes = [
                _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64,
                P(qpip_resource), _vp, P(_vp), P(_i64)]
        L.qpip_point_in_polygon_noalloc.argtypes = [
            _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp]
        L.qpip_batched_quadtree_point_in_polygon.argtypes = [
            _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
            _i64, P(qpip_resource), _vp, P(_vp), P(_vp), P(_vp), _vp, P(_i64)]
        L.qpip_batch_plan_create.argtypes = [
            _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
            _i64, _d, _vp, P(_vp)]
        L.qpip_batch_plan_launch.argtypes = [_vp, _vp]
        L.qpip_batch_plan_result.argtypes = [_vp, P(qpip_resource), _vp, P(_vp), P(_vp), P(_vp),
                                             _vp, P(_i64)]
        L.qpip_batch_plan_destroy.argtypes = [_vp]
        L.qpip_linestring_bounding_boxes.argtypes = [
            _vp, _i64, _vp, _vp, _i64, _i64, _d, P(qpip_resource), _vp,
            P(_vp), P(_vp), P(_vp), P(_vp), P(_i64)]
        L.qpip_quadtree_point_to_nearest_linestring.argtypes = [
            _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64,
            _vp, _i64, _vp, _vp, _i64, _i64, P(qpip_resource), _vp, P(_vp), P(_vp), P(_vp),
            P(_i64)]
        L.qpip_radix_sort_pairs.argtypes = [_vp, _vp, _i64, ctypes.c_int32, _vp]
        L.qpip_compact_nonnegative.argtypes = [_vp, _i64, _vp, _vp, P(_i64), _vp]
        L.qpip_ranker_num_params.argtypes = [_i64]
        L.qpip_ranker_num_params.restype = _i64
        L.qpip_ranker_features.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp]
        L.qpip_ranker_logits.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp]
        L.qpip_ranker_gradients.argtypes = [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]
        _f = ctypes.c_float
        L.qpip_ranker_adam.argtypes = [_vp, _vp, _vp, _vp, _vp, _i64, _f, _f, _f, _f, _i64, _vp]
        L.qpip_ranker_locate.argtypes = [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _i64,
                                         ctypes.c_int32, _vp, _vp, _vp, _vp]
        for name in (
            "qpip_linestring_bounding_boxes", "qpip_quadtree_point_to_nearest_linestring",
            "qpip_ranker_features", "qpip_ranker_logits", "qpip_ranker_gradients",
            "qpip_ranker_adam", "qpip_ranker_locate",
            "qpip_batched_quadtree_point_in_polygon", "qpip_batch_plan_create",
            "qpip_batch_plan_launch", "qpip_batch_plan_result", "qpip_batch_plan_destroy",
            "qpip_point_in_polygon", "qpip_pairwise_point_in_polygon",
            "qpip_point_in_polygon_noalloc",
            "qpip_quadtree_on_points", "qpip_quadtree_on_points_xy", "qpip_polygon_bounding_boxes",
            "qpip_join_quadtree_and_bounding_boxes", "qpip_quadtree_point_in_polygon",
            "qpip_quadtree_point_in_polygon_xy",
            "qpip_radix_sort_pairs", "qpip_compact_nonnegative",
        ):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def version() -> str:
    return lib().qpip_version().decode()


def profile_enable(on: bool = True) -> None:
    lib().qpip_profile_enable(1 if on else 0)


def profile_reset() -> None:
    lib().qpip_profile_reset()


def profile_report() -> dict:
    """{kernel name: (launches, total ms)} recorded while profiling was enabled."""
    import json

    L = lib()
    n = L.qpip_profile_report(None, 0)
    buf = ctypes.create_string_buffer(int(n) + 1)
    L.qpip_profile_report(buf, len(buf))
    return {k: (int(v[0]), float(v[1])) for k, v in json.loads(buf.value.decode()).items()}


def memory_info() -> tuple:
    """(live, cached) bytes of the library's pool on the current device."""
    live, cached = ctypes.c_size_t(), ctypes.c_size_t()
    lib().qpip_memory_info(ctypes.byref(live), ctypes.byref(cached))
    return int(live.value), int(cached.value)


def sort_pairs(keys: torch.Tensor, values: torch.Tensor, key_bits: int = 32) -> None:
    """Stable radix sort of (keys, values) in place: uint32/int32 tensors of
    equal length on one device, ordered by the low key_bits bits of the keys."""
    if keys.numel() != values.numel():
        raise ValueError("keys and values must have the same length")
    if keys.element_size() != 4 or values.element_size() != 4 or not keys.is_contiguous() \
            or not values.is_contiguous():
        raise ValueError("keys and values must be contiguous 32-bit tensors")
    check(lib().qpip_radix_sort_pairs(ctypes.c_void_p(keys.data_ptr()),
                                      ctypes.c_void_p(values.data_ptr()), keys.numel(),
                                      int(key_bits), stream_handle(keys.device)))


def compact_nonnegative(values: torch.Tensor, with_values: bool = True):
    """(indices, values) of the entries of an int32 tensor that are >= 0, in
    order; indices as int64, values as int32 (None without with_values)."""
    if values.dtype != torch.int32 or not values.is_contiguous():
        raise ValueError("values must be a contiguous int32 tensor")
    n = values.numel()
    idx = torch.empty(max(n, 1), dtype=torch.int32, device=values.device)
    val = torch.empty(max(n, 1), dtype=torch.int32, device=values.device) if with_values \
        else None
    cnt = ctypes.c_int64()
    check(lib().qpip_compact_nonnegative(
        ctypes.c_void_p(values.data_ptr()), n, ctypes.c_void_p(idx.data_ptr()),
        ctypes.c_void_p(val.data_ptr()) if val is not None else None, ctypes.byref(cnt),
        stream_handle(values.device)))
    k = int(cnt.value)
    return idx[:k].to(torch.int64), (val[:k] if val is not None else None)


def check(rc: int) -> None:
    if rc == QPIP_SUCCESS:
        return
    msg = lib().qpip_last_error_message().decode()
    if rc == QPIP_ERROR_OOM:
        raise MemoryError(msg)
    if rc == QPIP_ERROR_CAPACITY:
        raise CapacityError(msg)
    raise RuntimeError(msg)


class TorchResource:
    """A qpip_resource whose buffers are torch uint8 tensors on `device`.

    The library calls ``allocate`` for each output column; the tensors are kept
    here until ``take`` hands them to the caller (viewed with the right dtype).
    """

    def __init__(self, device: torch.device):
        self.device = device
        self._live: dict[int, torch.Tensor] = {}
        self._alloc_cb = _ALLOC(self._allocate)
        self._free_cb = _FREE(self._deallocate)
        self.struct = qpip_resource(self._alloc_cb, self._free_cb, None)

    def _allocate(self, ctx, nbytes, stream):
        try:
            t = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        except Exception:  # noqa: BLE001 - reported to the library as NULL (out of memory)
            return None
        p = t.data_ptr()
        self._live[p] = t
        return p

    def _deallocate(self, ctx, ptr, nbytes, stream):
        self._live.pop(ptr, None)

    def take(self, ptr, n: int, dtype: torch.dtype) -> torch.Tensor:
        if n == 0 or not ptr:
            return torch.empty(0, dtype=dtype, device=self.device)
        t = self._live.pop(ptr)
        return t.view(dtype)[:n]


def stream_handle(device: torch.device):
    """The HIP stream the library should use: torch's current stream on `device`."""
    if device.type != "cuda":
        return None
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
