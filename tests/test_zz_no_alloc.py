"""qpip_point_in_polygon_noalloc (the allocation-free variant of the brute-force
bitmask point-in-polygon) against qpip_point_in_polygon, both through the C ABI.

Two calls in a row on one stream, the second with fewer points, polygons, rings and
vertices than the first, so that it runs over the edge table the first call left in
the library's fixed scratch. Bar: bit-identical INT32 bitmasks; the library's pool
and torch's allocator hold the same bytes before and after the calls.
"""

import ctypes

import numpy as np
import pytest

from conftest import star_polygons

pytestmark = pytest.mark.gpu

MAX_VERTICES = 262144  # QPIP_NOALLOC_MAX_VERTICES
MAX_RINGS = 65536  # QPIP_NOALLOC_MAX_RINGS


@pytest.fixture(scope="module")
def env():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    import cuspatial_amd as cs
    from cuspatial_amd import _capi

    return torch, cs, _capi


def scene(seed, num_points, num_polys, kmax, holes):
    rng = np.random.default_rng(seed)
    part, ring, px, py = star_polygons(rng, num_polys, 1000.0, 20.0, 120.0, kmin=3, kmax=kmax,
                                       holes=holes)
    # points near the polygons, plus their vertices (on-edge cases)
    x = np.concatenate([rng.uniform(0, 1000, num_points - len(px)), px])
    y = np.concatenate([rng.uniform(0, 1000, num_points - len(py)), py])
    return x, y, part, ring, px, py


def call_noalloc(torch, L, x, y, part, ring, px, py, out, stream):
    return L.qpip_point_in_polygon_noalloc(
        x.data_ptr(), y.data_ptr(), x.numel(), y.numel(), part.data_ptr(), part.numel(),
        ring.data_ptr(), ring.numel(), px.data_ptr(), py.data_ptr(), px.numel(), py.numel(),
        ctypes.c_void_p(stream.cuda_stream), ctypes.c_void_p(out.data_ptr()))


def test_noalloc_matches_point_in_polygon_twice(env):
    torch, cs, _capi = env
    L = _capi.lib()
    dev = torch.device("cuda", 0)
    cases = [scene(11, 200_000, 31, 200, 0.5), scene(12, 5_000, 7, 12, 0.3)]
    assert len(cases[1][4]) < len(cases[0][4]) and len(cases[1][3]) < len(cases[0][3])
    stream = torch.cuda.Stream(device=dev)
    inputs, outs, wants = [], [], []
    for x, y, part, ring, px, py in cases:
        t = [torch.as_tensor(a, device=dev) for a in (x, y)]
        p = [torch.as_tensor(a, dtype=torch.int32, device=dev) for a in (part, ring)]
        v = [torch.as_tensor(a, device=dev) for a in (px, py)]
        wants.append(cs.point_in_polygon_bitmask(cs.Points(*t), cs.Polygons(*p, *v)))
        inputs.append((*t, *p, *v))
        outs.append(torch.full((len(x),), -7, dtype=torch.int32, device=dev))
    torch.cuda.synchronize()
    live0, cached0 = ctypes.c_size_t(), ctypes.c_size_t()
    L.qpip_memory_info(ctypes.byref(live0), ctypes.byref(cached0))
    torch_bytes0 = torch.cuda.memory_allocated(dev)
    print(f"noalloc: {len(cases[0][0])} then {len(cases[1][0])} points", flush=True)
    with torch.cuda.stream(stream):
        for args, out in zip(inputs, outs):
            rc = call_noalloc(torch, L, *args, out, stream)
            assert rc == 0, _capi.lib().qpip_last_error_message()
    stream.synchronize()
    live1, cached1 = ctypes.c_size_t(), ctypes.c_size_t()
    L.qpip_memory_info(ctypes.byref(live1), ctypes.byref(cached1))
    assert (live1.value, cached1.value) == (live0.value, cached0.value)
    assert torch.cuda.memory_allocated(dev) == torch_bytes0
    for out, want in zip(outs, wants):
        assert torch.equal(out, want.to(torch.int32))
    assert int((outs[0] != 0).sum()) > 1000  # the scene has hits


def test_noalloc_rejects_inputs_beyond_its_scratch(env):
    torch, cs, _capi = env
    L = _capi.lib()
    dev = torch.device("cuda", 0)
    x = torch.zeros(4, dtype=torch.float64, device=dev)
    out = torch.zeros(4, dtype=torch.int32, device=dev)
    part = torch.tensor([0, 1], dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    ring = torch.tensor([0, MAX_VERTICES + 1], dtype=torch.int32, device=dev)
    v = torch.zeros(MAX_VERTICES + 1, dtype=torch.float64, device=dev)
    rc = call_noalloc(torch, L, x, x, part, ring, v, v, out, stream)
    assert rc == _capi.QPIP_ERROR_CAPACITY
    assert b"262144" in L.qpip_last_error_message()
    ring = torch.zeros(MAX_RINGS + 2, dtype=torch.int32, device=dev)
    rc = call_noalloc(torch, L, x, x, part, ring, v[:4], v[:4], out, stream)
    assert rc == _capi.QPIP_ERROR_CAPACITY
    # a call within the limits still works after the rejections
    ring = torch.tensor([0, 4], dtype=torch.int32, device=dev)
    sq = torch.tensor([0.0, 1.0, 1.0, 0.0], dtype=torch.float64, device=dev)
    sy = torch.tensor([0.0, 0.0, 1.0, 0.0], dtype=torch.float64, device=dev)
    px = torch.tensor([0.75, 0.25, 5.0, 0.5], dtype=torch.float64, device=dev)
    py = torch.tensor([0.25, 0.75, 5.0, 0.1], dtype=torch.float64, device=dev)
    assert call_noalloc(torch, L, px, py, part, ring, sq, sy, out, stream) == 0
    want = cs.point_in_polygon_bitmask(cs.Points(px, py), cs.Polygons(part, ring, sq, sy))
    torch.cuda.synchronize()
    assert torch.equal(out, want.to(torch.int32))
    assert out.tolist() == [1, 0, 0, 1]
