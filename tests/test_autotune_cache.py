"""Autotuner cache logic on the CPU: a cached choice is applied without
timing; an unknown key or a stale candidate list would need the GPU."""
import json

from molmot_amd import autotune


def test_cached_choice_applied_without_timing(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "_device_key", lambda: "devX|gfx950|256")
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"devX|gfx950|256|k1": {"best": 128, "ms": {"64": 2.0, "128": 1.0, "256": 1.5}}}))
    t = autotune.Tuner(str(path))
    applied = []

    def run():
        raise AssertionError("a cached key must not be timed")

    best, ms = t.pick("k1", (64, 128, 256), applied.append, run)
    assert best == 128 and applied == [128] and ms["128"] == 1.0
    assert t.lookup("k1") == 128 and t.lookup("k2") is None


def test_no_cache_file(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "_device_key", lambda: "devX")
    t = autotune.Tuner(None)
    assert t.lookup("anything") is None
    t._save()   # no path: nothing written
    assert not list(tmp_path.iterdir())
