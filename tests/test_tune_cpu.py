"""CPU checks of the tuning file format (segclip.tune): save/load round trip,
validation of names and ranges, and the committed tuning file loads."""
import json
import os

import pytest

from segclip import tune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_save_load_round_trip(tmp_path):
    p = str(tmp_path / "t.json")
    rec = tune.save(p, {"a": {"settings": {"stab_tiles": 4}, "workload": "w",
                              "timings_ms": {"stab_tiles": {1: 5.0, 2: 4.5, 4: 4.25}}},
                        "b": {"settings": {"stab_tiles": 1}}}, "dev")
    assert rec["workloads"]["a"]["settings"] == {"stab_tiles": 4}
    assert tune.load(p, "a") == {"stab_tiles": 4}
    assert tune.load(p, "b") == {"stab_tiles": 1}
    assert tune.load(p, "missing") == {}
    with open(p) as f:
        assert json.load(f)["workloads"]["a"]["timings_ms"]["stab_tiles"]["4"] == 4.25


@pytest.mark.parametrize("bad", [{"stab_tiles": 0}, {"stab_tiles": 17}, {"nope": 2}])
def test_invalid_settings_refused(tmp_path, bad):
    with pytest.raises(ValueError):
        tune.check(bad)
    p = str(tmp_path / "bad.json")
    with open(p, "w") as f:
        json.dump({"workloads": {"x": {"settings": bad}}}, f)
    with pytest.raises(ValueError):
        tune.load(p, "x")


def test_defaults_are_candidates():
    for k, v in tune.DEFAULTS.items():
        assert v in tune.SETTINGS[k]
        tune.check({k: v})


def test_committed_tuning_file_loads():
    p = os.path.join(ROOT, "segment-tree-based-clipping_amd", "tuning.json")
    if not os.path.exists(p):
        pytest.skip("no tuning file committed")
    with open(p) as f:
        names = list(json.load(f)["workloads"])
    assert "plain" in names
    for n in names:
        assert set(tune.load(p, n)) <= set(tune.SETTINGS)
