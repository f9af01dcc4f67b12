"""The autotuner's tuning table (host only): shape buckets, the JSON round trip,
version and parameter checks, and scoped application of a configuration."""

import json

import pytest


@pytest.fixture(scope="module")
def autotune():
    from cuspatial_amd import _build, _capi, autotune

    import os

    if not os.path.exists(_capi.LIB_PATH):
        _build.build()
    return autotune


def test_shape_bucket_rounds_log2():
    from cuspatial_amd.autotune import shape_bucket

    assert shape_bucket(4096, 4096 * 145_000, 1_078_129, 141_228_680 - 1_078_129) == \
        "jobs12_points17_edges7"
    assert shape_bucket(1, 1000, 10, 40) == "jobs0_points10_edges2"
    # empty shapes stay defined
    assert shape_bucket(0, 0, 0, 0) == "jobs0_points0_edges0"


def test_table_round_trip(tmp_path, autotune):
    t = autotune.TuningTable(None)
    res = {"config": {"pip_pairs_shift": 20, "pip_waves_per_block": 4},
           "baseline_ms": 2.0, "best_ms": 1.5}
    t.record("jobs3_points10_edges5", res, "test")
    path = str(tmp_path / "tuning.json")
    t.save(path)
    u = autotune.TuningTable(path)
    assert u.lookup("jobs3_points10_edges5") == res["config"]
    assert u.lookup("jobs4_points10_edges5") is None
    with open(path) as f:
        data = json.load(f)
    assert data["version"] == 1


def test_table_rejects_other_versions_and_unknown_parameters(tmp_path, autotune):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"version": 2, "buckets": {}}))
    with pytest.raises(ValueError, match="version"):
        autotune.TuningTable(str(p))
    p.write_text(json.dumps({"version": 1, "buckets": {"b": {"config": {"nope": 1}}}}))
    with pytest.raises(ValueError, match="unknown tuning parameters"):
        autotune.TuningTable(str(p))


def test_applied_restores_the_previous_values(autotune):
    before = autotune.current()
    with autotune.applied({"pip_pairs_shift": 12, "pip_stage_edges": 128}):
        assert autotune.get("pip_pairs_shift") == 12
        assert autotune.get("pip_stage_edges") == 128
    assert autotune.current() == before
    with pytest.raises(RuntimeError):
        with autotune.applied({"pip_pairs_shift": 99}):
            pass
    assert autotune.current() == before


def test_package_table_loads(autotune):
    """The committed table (if any) has the current format and known parameters."""
    import os

    if os.path.exists(autotune.DEFAULT_TABLE):
        t = autotune.TuningTable(autotune.DEFAULT_TABLE)
        for b in t.buckets:
            assert b.startswith("jobs")
