"""CPU tests of the autotuner's cache logic (no GPU: the ensemble is a stub
that records calls)."""
import json

import numpy as np

from palinstrophy_amd import autotune


class StubEnsemble:
    precision = 4

    def __init__(self):
        self.cls = [[256, 4, 0, 1], [512, 2, 0, 1]]
        self.tuned = 0
        self.applied = []

    def classes(self):
        return [tuple(c) for c in self.cls]

    def tune(self, reps):
        self.tuned += 1
        ms = np.full((2, 2, 4), -1.0, dtype=np.float32)
        ms[0, 0] = [3.0, 2.0, 1.0, 4.0]
        ms[0, 1] = [5.0, 6.0, 7.0, 4.5]
        ms[1, 0] = [1.0, 2.0, 3.0, 4.0]
        ms[1, 1] = [2.0, 1.0, 3.0, 4.0]
        self.cls[0][2:] = [2, 3]
        self.cls[1][2:] = [0, 1]
        return ms

    def set_variants(self, N, vr, vc):
        self.applied.append((N, vr, vc))


def test_cache_roundtrip_and_reuse(tmp_path, monkeypatch):
    path = str(tmp_path / "sub" / "tune.json")
    monkeypatch.setattr(autotune, "device_id", lambda: "dev|gfx950")
    monkeypatch.setattr(autotune, "library_id", lambda: "123-456")
    e = StubEnsemble()
    r1 = autotune.tune_ensemble(e, reps=2, path=path)
    assert e.tuned == 1 and e.applied == []
    assert r1[256]["vrow"] == 2 and r1[256]["vcol"] == 3 and not r1[256]["cached"]
    data = json.load(open(path))
    key = autotune.entry_key("dev|gfx950", "123-456", 256, 4, 4)
    assert data[key]["vrow"] == 2 and data[key]["ms_row"][2] == 1.0
    e2 = StubEnsemble()
    r2 = autotune.tune_ensemble(e2, reps=2, path=path)
    assert e2.tuned == 0
    assert sorted(e2.applied) == [(256, 2, 3), (512, 0, 1)]
    assert all(v["cached"] for v in r2.values())


def test_cache_keyed_by_library_build(tmp_path, monkeypatch):
    path = str(tmp_path / "tune.json")
    monkeypatch.setattr(autotune, "device_id", lambda: "dev")
    monkeypatch.setattr(autotune, "library_id", lambda: "build-a")
    autotune.tune_ensemble(StubEnsemble(), path=path)
    monkeypatch.setattr(autotune, "library_id", lambda: "build-b")
    e = StubEnsemble()
    autotune.tune_ensemble(e, path=path)
    assert e.tuned == 1  # a rebuilt library is tuned again


def test_unreadable_cache_is_ignored(tmp_path):
    p = tmp_path / "tune.json"
    p.write_text("{not json")
    assert autotune.load_cache(str(p)) == {}
    assert autotune.load_cache(str(tmp_path / "absent.json")) == {}
