"""Autotuner table: record, look up and apply (CPU; the GPU timing itself is
covered by test_gpu_parity.test_encode_tiles_exact; the pipeline tuner is switched off,
see test_pipeline_tuner_is_switched_off)."""
import importlib.util
import json
from pathlib import Path

import pytest

TUNE = Path(__file__).resolve().parents[1] / "fherma-bytecrt_amd" / "tune.py"


@pytest.fixture(scope="module")
def tune():
    spec = importlib.util.spec_from_file_location("bcrt_tune_cpu", TUNE)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class FakeEngine:
    def __init__(self, N, W, max_batch=128):
        self.N, self.W, self.max_batch = N, W, max_batch
        self.calls = []

    def set_pipeline(self, chunk, streams):
        self.calls.append((chunk, streams))

    def set_encode_tile(self, ts_log):
        self.calls.append(("tile", ts_log))

    def set_fusion(self, min_tiles):
        self.calls.append(("fuse", min_tiles))

    def set_kernel_flags(self, flags):
        self.calls.append(("flags", flags))


def test_record_lookup_apply(tune, tmp_path):
    path = tmp_path / "table.json"
    res = {"best": {"chunk": 32, "streams": 4, "ms": 17.5}, "rows": []}
    tune.record(65536, 1747, 128, res, device="test", path=path)
    table = tune.load_table(path)
    assert tune.lookup(65536, 1747, 128, table) == (32, 4)
    assert tune.lookup(65536, 1747, 64, table) is None       # batch is part of the key
    assert tune.lookup(32768, 868, 128, table) is None
    e = FakeEngine(65536, 1747)
    assert tune.apply(e, 128, table) and e.calls == [(32, 4)]
    assert not tune.apply(e, 1, table) and e.calls == [(32, 4)]
    # a second record keeps the first entry
    tune.record(32768, 868, 12, {"best": {"chunk": 8, "streams": 2, "ms": 1.0}, "rows": []}, "test", path)
    table = json.loads(path.read_text())
    assert set(table) == {"N65536_W1747_B128", "N32768_W868_B12"}


def test_encode_tile_entries(tune, tmp_path):
    path = tmp_path / "table.json"
    enc = {"best": {"ts_log": 4, "tile": "8 r x 16 s", "encode_ms": 4.0}, "rows": []}
    # an encode-only entry, then the pipeline added to it: both kept
    tune.record(65536, 1747, 128, None, device="test", path=path, encode=enc)
    table = tune.load_table(path)
    assert tune.lookup(65536, 1747, 128, table) is None
    assert tune.lookup_encode_tile(65536, 1747, 128, table) == 4
    e = FakeEngine(65536, 1747)
    assert tune.apply(e, 128, table) and e.calls == [("tile", 4)]
    tune.record(65536, 1747, 128, {"best": {"chunk": 32, "streams": 2, "ms": 17.0}, "rows": []}, "test", path)
    table = tune.load_table(path)
    assert tune.lookup(65536, 1747, 128, table) == (32, 2)
    assert tune.lookup_encode_tile(65536, 1747, 128, table) == 4
    e = FakeEngine(65536, 1747)
    assert tune.apply(e, 128, table) and e.calls == [(32, 2), ("tile", 4)]


def test_fusion_entries(tune, tmp_path):
    path = tmp_path / "table.json"
    fus = {"best": {"fuse_min_tiles": tune.FUSE_NEVER, "structure": "separate", "encode_fwd_ms": 0.01}, "rows": []}
    tune.record(16384, 436, 1, None, device="test", path=path, fusion=fus)
    table = tune.load_table(path)
    assert tune.lookup_fusion(16384, 436, 1, table) == 0xFFFFFFFF
    assert tune.lookup(16384, 436, 1, table) is None
    e = FakeEngine(16384, 436)
    assert tune.apply(e, 1, table) and e.calls == [("fuse", 0xFFFFFFFF)]


def test_encode_tile_candidates(tune):
    class G:
        def __init__(self, N1):
            self.N1 = N1
    assert tune.encode_tile_candidates(G(256)) == tuple(range(8))   # N = 65536: 128 r x 1 s .. 1 r x 128 s
    assert tune.encode_tile_candidates(G(32)) == tuple(range(6))    # N = 8192: at most 32 s
    assert tune.encode_tile_candidates(G(16)) == tuple(range(5))


def test_table_from_environment(tune, tmp_path, monkeypatch):
    path = tmp_path / "env.json"
    path.write_text(json.dumps({"N8192_W220_B4": {"chunk": 2, "streams": 2, "ms": 0.1}}))
    monkeypatch.setenv("BCRT_PIPELINE_TABLE", str(path))
    assert tune.lookup(8192, 220, 4) == (2, 2)


def test_shipped_table_is_valid(tune):
    table = tune.load_table(tune.TABLE)
    for key, e in table.items():
        assert key.startswith("N") and "_W" in key and "_B" in key
        if "chunk" in e:
            assert 1 <= e["chunk"] and 1 <= e["streams"] <= 16
        if "enc_ts_log" in e:
            assert 0 <= e["enc_ts_log"] <= 7
        if "fuse_min_tiles" in e:
            assert 0 <= e["fuse_min_tiles"] <= 0xFFFFFFFF


def test_pipeline_tuner_is_switched_off(tune):
    """The pipeline tuner raises instead of capturing graphs (its capture after
    an eager run crashed the host process twice, cause not established)."""
    with pytest.raises(RuntimeError, match="switched off"):
        tune.tune_pipeline(None, None, None, None, 1)


def test_kernel_flag_entries(tune, tmp_path):
    """Kernel flags are kept beside the other choices of an entry and applied
    with them; an entry without them leaves the engine's flags alone."""
    path = tmp_path / "table.json"
    fus = {"best": {"fuse_min_tiles": 0, "structure": "fused", "encode_fwd_ms": 5.8}, "rows": []}
    flg = {"best": {"kernel_flags": 2, "product_ms": 14.6}, "rows": []}
    tune.record(65536, 1747, 128, None, device="test", path=path, fusion=fus)
    tune.record(65536, 1747, 128, None, device="test", path=path, flags=flg)
    table = tune.load_table(path)
    assert tune.lookup_kernel_flags(65536, 1747, 128, table) == 2
    assert tune.lookup_fusion(65536, 1747, 128, table) == 0
    e = FakeEngine(65536, 1747)
    assert tune.apply(e, 128, table) and ("flags", 2) in e.calls and ("fuse", 0) in e.calls
    assert tune.lookup_kernel_flags(32768, 868, 128, table) is None
    assert tune.KERNEL_FLAG_CANDIDATES == tuple(range(32))
