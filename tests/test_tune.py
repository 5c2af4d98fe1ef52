"""The tuner's bookkeeping on the CPU: choices are stored per device key and
workload, other entries survive a save, an unknown device or workload loads
as None, and a damaged file loads as None instead of raising."""
import json

import pytest

from dyncolor_amd import tune


@pytest.fixture
def key(monkeypatch):
    k = {"v": "devA|gfx950|256"}
    monkeypatch.setattr(tune, "_key", lambda: k["v"])
    return k


def test_save_load_merge(tmp_path, key):
    p = str(tmp_path / "tuning.json")
    assert tune.load(p, "c2") is None
    tune.save(p, "c2", {"driver": 0, "light_c": 4096}, [{"param": "light_c"}], {"batch": 10})
    tune.save(p, "c4", {"driver": 2}, [])
    key["v"] = "devB|gfx950|256"
    tune.save(p, "c2", {"driver": 0, "light_c": 65536}, [])
    assert tune.load(p, "c2") == {"driver": 0, "light_c": 65536}
    assert tune.load(p, "c4") is None
    key["v"] = "devA|gfx950|256"
    assert tune.load(p, "c2") == {"driver": 0, "light_c": 4096}
    assert tune.load(p, "c4") == {"driver": 2}
    data = json.load(open(p))
    assert data["devA|gfx950|256"]["c2"]["batch"] == 10
    assert set(data) == {"devA|gfx950|256", "devB|gfx950|256"}


def test_load_damaged_file(tmp_path, key):
    p = tmp_path / "tuning.json"
    p.write_text("{not json")
    assert tune.load(str(p), "c2") is None


def test_candidates_are_valid_tunings():
    # every candidate value is inside the ranges dc_set_tuning accepts
    for name, vals in tune.CANDIDATES.items():
        assert vals, name
        if name == "driver":
            continue
        for v in vals:
            if name in ("loop_blocks",):
                assert 1 <= v <= 4096
            elif name == "light_blocks":
                assert 0 <= v <= 4096
            else:
                assert v >= 0


def test_descent_takes_only_clear_wins():
    # a fake trial time: light_c = 1024 is 20 % faster, list_c = 16384 1 %
    # faster (inside the margin); the descent takes the first and not the
    # second, tries each candidate at most once (the old default again after
    # 1024 is taken) and keeps the choice in state
    def run(p, j):
        t = 1.0
        if p["light_c"] == 1024:
            t *= 0.8
        if p["list_c"] == 16384:
            t *= 0.99
        return t
    state = {}
    cands = {k: tune.CANDIDATES[k] for k in ("light_c", "list_c")}
    best, sweep = tune._descend(run, 2, 0.03, None, None, cands, state)
    assert best["light_c"] == 1024 and best["list_c"] != 16384
    assert state["best"] == best
    tried = [e["value"] for e in sweep if e["param"] == "light_c"]
    assert sorted(tried) == sorted(set(tried)) and 1024 in tried
    assert [e for e in sweep if e.get("taken")] == [e for e in sweep if e["value"] == 1024]
