"""The autotuner's persisted choices: file naming per device, loading, and
which parameters reach SPD and LDL^T plans (no GPU needed)."""
import json

from opt_solver_amd import _lib, autotune as at


class FakePlan:
    def __init__(self):
        self.set = {}

    def tune(self, key, value):
        self.set[key] = value


def test_device_key_and_path():
    assert at.device_key("AMD Instinct MI355X (gfx950)") == "AMD_Instinct_MI355X_gfx950"
    assert at.tuned_path("AMD Radeon Graphics (gfx950)").endswith("tuned/AMD_Radeon_Graphics_gfx950.json")


def test_every_knob_is_a_library_key():
    for bucket, name, cands, default, used in at.KNOBS:
        assert hasattr(_lib, "TUNE_" + name), name
        assert default in cands and len(set(cands)) == len(cands)
        assert bucket and used


def test_load_and_apply(tmp_path, monkeypatch):
    monkeypatch.setattr(at, "TUNED_DIR", str(tmp_path))
    assert at.load_tuned("Some GPU") is None
    knobs = {"LDLT_BIG_MIN": 384, "LDLT_UPDATE_WG": 2048, "PANEL_OUTER": 256, "CB_SEG": 2048}
    with open(at.tuned_path("Some GPU"), "w") as fh:
        json.dump({"device": "Some GPU", "knobs": knobs}, fh)
    t = at.load_tuned("Some GPU")
    assert t["knobs"] == knobs
    # a file of another device is not taken
    with open(at.tuned_path("Other GPU"), "w") as fh:
        json.dump({"device": "Some GPU", "knobs": knobs}, fh)
    assert at.load_tuned("Other GPU") is None
    spd, ldlt = FakePlan(), FakePlan()
    assert at.apply_tuned(spd, t, posdef=True) == {"PANEL_OUTER": 256, "CB_SEG": 2048}
    assert at.apply_tuned(ldlt, t, posdef=False) == {"LDLT_BIG_MIN": 384, "LDLT_UPDATE_WG": 2048}
    assert spd.set == {_lib.TUNE_PANEL_OUTER: 256, _lib.TUNE_CB_SEG: 2048}
    assert ldlt.set == {_lib.TUNE_LDLT_BIG_MIN: 384, _lib.TUNE_LDLT_UPDATE_WG: 2048}
