"""The chunk-tuning table and its file (CPU only: host-side calls of the C ABI, no device)."""
import json

import pytest

from thewalrus_amd import _native, tune


@pytest.fixture
def restore_table():
    before = tune.current()
    yield
    tune.apply([{"kind": k, "T": T, "loop": lp, "blocks_per_request": v} for (k, T, lp), v in before.items()])


def test_set_get_round_trip(restore_table):
    L = _native.load()
    assert L.wh_tune_set(0, 5, 1, 2048) == 0
    assert L.wh_tune_get(0, 5, 1) == 2048
    assert L.wh_tune_set(1, 0, 0, 64) == 0
    assert tune.current()[(1, 0, 0)] == 64


@pytest.mark.parametrize("kind, T, loop, blocks", [(0, 0, 0, 1), (0, 7, 0, 1), (1, 1, 0, 1), (2, 0, 0, 1),
                                                   (0, 3, 2, 1), (0, 3, 0, -1)])
def test_out_of_range_rejected(kind, T, loop, blocks, restore_table):
    L = _native.load()
    assert L.wh_tune_set(kind, T, loop, blocks) != 0
    assert b"out of range" in L.wh_last_error()
    with pytest.raises(ValueError):
        tune.apply([{"kind": kind, "T": T, "loop": loop, "blocks_per_request": blocks}])


def test_file_applied(tmp_path, restore_table):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"entries": [{"kind": 0, "T": 4, "loop": 0, "blocks_per_request": 512},
                                         {"kind": 1, "T": 0, "loop": 1, "blocks_per_request": 128}]}))
    entries = tune.load(str(p))
    assert len(entries) == 2
    t = tune.current()
    assert t[(0, 4, 0)] == 512 and t[(1, 0, 1)] == 128
    assert tune.load(str(tmp_path / "missing.json")) == []


def test_environment_switch(monkeypatch):
    monkeypatch.setenv("WH_TUNING", "none")
    assert tune.table_path() is None
    monkeypatch.setenv("WH_TUNING", "/x/y.json")
    assert tune.table_path() == "/x/y.json"
    monkeypatch.delenv("WH_TUNING")
    assert tune.table_path() == tune.DEFAULT_PATH


def test_choose_keeps_default_within_noise():
    assert tune.choose(0, {256: 9.0, 1024: 9.5, 4096: 9.8}) == 256
    assert tune.choose(0, {256: 9.45, 1024: 9.5}) == 1024  # 0.5% is noise
    assert tune.choose(1, {256: 4.0, 2048: 3.9}) == 2048
    assert tune.choose(1, {256: 4.0, 2048: 3.99}) == 256


def test_committed_table_is_valid():
    import os

    assert os.path.exists(tune.DEFAULT_PATH)
    with open(tune.DEFAULT_PATH) as f:
        doc = json.load(f)
    keys = {(e["kind"], e["T"], e["loop"]) for e in doc["entries"]}
    assert keys == set(tune.current().keys())
    for e in doc["entries"]:
        assert e["blocks_per_request"] == tune.choose(e["kind"], {int(k): v for k, v in e["ms"].items()})
        if "waves_ms" in e:
            assert e["waves"] == tune.choose_waves({int(k): v for k, v in e["waves_ms"].items()})
        if "tile_ms" in e:
            assert e["tile"] == tune.choose_waves({int(k): v for k, v in e["tile_ms"].items()})


def test_waves_round_trip_and_range(restore_table):
    L = _native.load()
    before = tune.current_waves()
    try:
        assert L.wh_tune_set_waves(0, 5, 1, 2) == 0 and L.wh_tune_get_waves(0, 5, 1) == 2
        assert L.wh_tune_set_waves(2, 2, 0, 4) == 0 and L.wh_tune_get_waves(2, 2, 0) == 4
        for args in ((1, 0, 0, 1), (0, 7, 0, 1), (0, 3, 0, 5), (2, 4, 0, 1), (0, 3, 0, -1)):
            assert L.wh_tune_set_waves(*args) != 0
        assert L.wh_tune_get_waves(1, 0, 0) == -1
    finally:
        for (k, T, lp), w in before.items():
            L.wh_tune_set_waves(k, T, lp, w)


def test_choose_waves_keeps_default_within_noise():
    assert tune.choose_waves({0: 10.0, 1: 9.95, 2: 12.0}) == 0
    assert tune.choose_waves({0: 10.0, 2: 9.5}) == 2


def test_tile_round_trip_and_range(restore_table):
    L = _native.load()
    before = tune.current_tiles()
    try:
        assert L.wh_tune_set_tile(0, 3, 1, 4) == 0 and L.wh_tune_get_tile(0, 3, 1) == 4
        assert L.wh_tune_set_tile(0, 5, 0, 8) == 0 and L.wh_tune_get_tile(0, 5, 0) == 8
        assert L.wh_tune_set_tile(2, 1, 0, 2) == 0 and L.wh_tune_get_tile(2, 1, 0) == 2
        assert L.wh_tune_set_tile(2, 1, 0, 3) == 0 and L.wh_tune_get_tile(2, 1, 0) == 3
        assert L.wh_tune_set_tile(2, 3, 1, 2) == 0 and L.wh_tune_get_tile(2, 3, 1) == 2
        assert L.wh_tune_set_tile(0, 3, 1, 0) == 0 and L.wh_tune_get_tile(0, 3, 1) == 0
        # a side not above the class's own (repeated rows: by the 2E it holds), past the kernels'
        # largest, a general-kernel class
        for args in ((0, 3, 0, 3), (0, 3, 0, 2), (0, 5, 0, 9), (2, 3, 0, 1), (2, 3, 0, 3), (2, 2, 0, 3),
                     (2, 1, 0, 4), (1, 0, 0, 2)):
            assert L.wh_tune_set_tile(*args) != 0
        assert L.wh_tune_get_tile(1, 0, 0) == -1
    finally:
        for (k, T, lp), t in before.items():
            L.wh_tune_set_tile(k, T, lp, t)
