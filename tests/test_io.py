"""GeoArrow input/output and the versioned index file (CPU; host arrays only)."""

import numpy as np
import pyarrow as pa
import pytest

from conftest import star_polygons


def np_(t):
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def io():
    from cuspatial_amd import io

    return io


def test_points_layout_and_round_trip(io, tmp_path):
    """A GeoArrow point column is fixed_size_list<double>[2] holding x0, y0, x1, y1..."""
    from cuspatial_amd import Points

    xy = np.array([[0.5, 1.5], [2.0, -3.0], [1e300, -0.0]])
    P = Points.from_points_xy(xy.reshape(-1), device="cpu")
    arr = io.points_to_arrow(P)
    assert arr.type == pa.list_(pa.float64(), 2)
    assert arr.to_pylist() == xy.tolist()
    for fmt in ("ipc", "parquet"):
        path = tmp_path / f"p.{fmt}"
        getattr(io, f"write_{fmt}")(path, pts=P)
        back = getattr(io, f"read_{fmt}")(path, device="cpu")["pts"]
        assert np.array_equal(np_(back.xy), xy.reshape(-1))
        assert np.signbit(np_(back.y)[2])  # -0.0 survives


def test_points_other_encodings(io):
    """Separated struct<x, y> coordinates, the reference's list<double> points,
    its dense union with only points, sliced and chunked arrays."""
    xy = np.arange(20, dtype=np.float64).reshape(-1, 2)
    struct = pa.StructArray.from_arrays([pa.array(xy[:, 0]), pa.array(xy[:, 1])], ["x", "y"])
    lists = pa.array(xy.tolist(), pa.list_(pa.float64()))
    union = pa.UnionArray.from_dense(pa.array([0] * 10, pa.int8()), pa.array(range(10), pa.int32()),
                                     [lists], ["points"])
    for a in (struct, lists, union, pa.chunked_array([lists[:4], lists[4:]])):
        got = io.points_from_arrow(a, device="cpu")
        assert np.array_equal(np.stack([np_(got.x), np_(got.y)], 1), xy)
    got = io.points_from_arrow(io.points_to_arrow(
        io.points_from_arrow(lists, device="cpu")).slice(3, 4), device="cpu")
    assert np.array_equal(np_(got.xy), xy[3:7].reshape(-1))
    with pytest.raises(ValueError):
        io.points_from_arrow(pa.array([[1.0, 2.0, 3.0]], pa.list_(pa.float64())), device="cpu")


def test_polygons_layout_and_round_trip(io, tmp_path):
    """polygon -> list<list<coord>>: one entry per polygon, its rings, their
    vertices; the offsets are the part / ring offsets."""
    from cuspatial_amd import Polygons

    rng = np.random.default_rng(3)
    part, ring, vx, vy = star_polygons(rng, 25, 100.0, 2.0, 10.0, holes=0.4)
    P = Polygons(part, ring, vx, vy, device="cpu")
    arr = io.polygons_to_arrow(P)
    assert arr.type == pa.list_(pa.list_(pa.list_(pa.float64(), 2)))
    first = arr[0].as_py()
    assert first[0] == [[vx[i], vy[i]] for i in range(ring[0], ring[1])]
    assert len(arr[0].as_py()) == part[1] - part[0]
    for fmt in ("ipc", "parquet"):
        path = tmp_path / f"q.{fmt}"
        getattr(io, f"write_{fmt}")(path, polys=P)
        back = getattr(io, f"read_{fmt}")(path, device="cpu")["polys"]
        assert np.array_equal(np_(back.part_offset), part)
        assert np.array_equal(np_(back.ring_offset), ring)
        assert np.array_equal(np_(back.x), vx) and np.array_equal(np_(back.y), vy)
        assert back.geometry_offset is None


def test_multipolygons_and_reference_layout(io, tmp_path):
    """multipolygon -> list<list<list<coord>>> keeps geometry offsets; the
    reference's list<list<list<list<double>>>> polygons read the same way."""
    from cuspatial_amd import Polygons

    rng = np.random.default_rng(4)
    part, ring, vx, vy = star_polygons(rng, 12, 100.0, 2.0, 10.0, holes=0.5)
    geom = np.array([0, 2, 3, 7, 12], np.int32)
    P = Polygons(part, ring, vx, vy, geometry_offset=geom, device="cpu")
    path = tmp_path / "m.arrow"
    io.write_ipc(path, polys=P)
    back = io.read_ipc(path, device="cpu")["polys"]
    assert np.array_equal(np_(back.geometry_offset), geom)
    assert np.array_equal(np_(back.part_offset), part)
    nested = io.polygons_to_arrow(P).to_pylist()
    ref_layout = pa.array(nested, pa.list_(pa.list_(pa.list_(pa.list_(pa.float64())))))
    again = io.polygons_from_arrow(ref_layout, device="cpu")
    assert np.array_equal(np_(again.ring_offset), ring)
    assert np.array_equal(np_(again.x), vx)
    # a slice rebases the offsets
    sl = io.polygons_from_arrow(io.polygons_to_arrow(Polygons(part, ring, vx, vy,
                                                              device="cpu")).slice(5, 3),
                                device="cpu")
    assert np_(sl.part_offset)[0] == 0 and np_(sl.ring_offset)[0] == 0
    assert np.array_equal(np_(sl.x), vx[ring[part[5]]:ring[part[8]]])


def test_index_file_round_trip_and_version(io, tmp_path):
    arrays = {"key": np.arange(7, dtype=np.uint32), "level": np.array([0, 1, 1], np.uint8),
              "is_internal_node": np.array([True, False, False]),
              "weights": np.ones((3, 4), np.float32), "x": np.linspace(0, 1, 5)}
    path = tmp_path / "idx.arrow"
    io.save_index(path, arrays, {"max_depth": 7, "scale": 7.8125})
    got, meta = io.load_index(path)
    assert meta == {"max_depth": 7, "scale": 7.8125}
    for k, v in arrays.items():
        assert got[k].dtype == v.dtype and got[k].shape == v.shape and np.array_equal(got[k], v)
    # a newer version is refused, another file is refused
    t = pa.ipc.open_file(pa.memory_map(str(path))).read_all()
    md = dict(t.schema.metadata)
    md[b"version"] = str(io.INDEX_VERSION + 1).encode()
    newer = tmp_path / "newer.arrow"
    with pa.OSFile(str(newer), "wb") as s, pa.ipc.new_file(s, t.schema.with_metadata(md)) as w:
        w.write_table(t.replace_schema_metadata(md))
    with pytest.raises(ValueError, match="version"):
        io.load_index(newer)
    other = tmp_path / "other.arrow"
    io.write_ipc(other, pts=__import__("cuspatial_amd").Points([0.0], [1.0], device="cpu"))
    with pytest.raises(ValueError, match="not a"):
        io.load_index(other)
