"""The launch-shape tuner's host logic and the recorded shapes (CPU only): coordinate
descent keeps the current shape unless a candidate is faster by more than the margin,
and every recorded lattice names every hlbl_tuning field with an accepted value."""
import ctypes

import pytest

import hlbl_amd as H
from hlbl_amd import tune as T

# the ranges hlbl_contractor_set_tuning accepts
RANGES = {"pass1_batch": (1, 1 << 20), "axes_wg_per_cu": (1, 16), "near_wg_per_cu": (1, 8),
          "pass2_wg_per_cu": (1, 4), "fp64_field_wg_per_cu": (1, 16), "fp64_pass2_wg_per_cu": (1, 8)}


def defaults():
    t = H.Tuning()
    H.check(H.lib().hlbl_tuning_defaults(ctypes.byref(t)))
    return t.as_dict()


def test_recorded_shapes_are_valid():
    import json
    with open(H.TUNING_PATH) as f:
        table = json.load(f)
    fields = {n for n, _ in H.Tuning._fields_}
    assert "48x24x24x24" in table["lattices"]
    for key, entry in table["lattices"].items():
        picks = entry["picks"]
        assert set(picks) == fields, key
        assert set(entry["defaults"]) == fields, key
        for n, (lo, hi) in RANGES.items():
            assert lo <= picks[n] <= hi, (key, n)
        assert picks["field_waves"] in (4, 8)
        assert 1 <= picks["field_wg_per_cu"] <= (2 if picks["field_waves"] == 8 else 3)
        assert H.tuned_shapes(tuple(int(d) for d in key.split("x"))) == picks
    assert H.tuned_shapes((7, 7, 7, 7)) is None


class FakeContractor:
    def __init__(self):
        self.t = defaults()

    def tuning(self):
        return dict(self.t)

    def set_tuning(self, **fields):
        assert set(fields) == set(self.t)
        self.t = dict(fields)


def test_tuner_keeps_the_current_shape_within_the_margin(monkeypatch):
    """pass1_batch 64 is 0.5% faster than the default 32: kept at 32 (margin 1%);
    pass2_wg_per_cu 2 is 5% faster than 1: taken; the field shape (4, 3) is the
    fastest: taken together; every other knob's times are flat: defaults kept."""
    times = {("pass1_batch", 64): 9.95, ("pass2_wg_per_cu", 2): 9.5,
             (("field_waves", "field_wg_per_cu"), (4, 3)): 8.0}
    con = FakeContractor()
    seen = []

    kinds = {"pass1_batch": "pass1", "pass2_wg_per_cu": "pass2", ("field_waves", "field_wg_per_cu"): "fields"}

    def fake_measure(c, descs, xunit, kind, reps):
        # the stage a knob affects reads only that knob's shape
        assert c is con
        t = c.tuning()
        for (knob, val), ms in times.items():
            names = knob if isinstance(knob, tuple) else (knob,)
            vals = val if isinstance(val, tuple) else (val,)
            if kinds[knob] == kind and tuple(t[n] for n in names) == vals:
                seen.append((knob, val))
                return ms
        return 10.0

    monkeypatch.setattr(T, "measure", fake_measure)
    monkeypatch.setattr(T, "log", lambda *a: None)
    base = defaults()
    picks, timings = T.tune(con, descs=[], xunit=0.05, reps=1, margin=0.01)
    assert picks["pass1_batch"] == base["pass1_batch"] == 32
    assert picks["pass2_wg_per_cu"] == 2
    assert (picks["field_waves"], picks["field_wg_per_cu"]) == (4, 3)
    for n in ("axes_wg_per_cu", "near_wg_per_cu", "fp64_field_wg_per_cu", "fp64_pass2_wg_per_cu"):
        assert picks[n] == base[n]
    assert con.tuning() == picks
    # every candidate of every knob was timed, in the tuner's order
    assert [k for k, _ in timings.items()] == ["+".join(k) if isinstance(k, tuple) else k
                                               for k, _, _ in T.CANDIDATES]
    for knob, cands, _ in T.CANDIDATES:
        key = "+".join(knob) if isinstance(knob, tuple) else knob
        assert sorted(timings[key]) == sorted(str(c) for c in cands)
    assert ("pass1_batch", 64) in seen


def test_tuner_candidates_are_accepted_values():
    for knob, cands, kind in T.CANDIDATES:
        assert kind in ("pass1", "pass2", "fields", "plain")
        for c in cands:
            vals = dict(zip(knob, c)) if isinstance(knob, tuple) else {knob: c}
            for n, v in vals.items():
                if n in RANGES:
                    assert RANGES[n][0] <= v <= RANGES[n][1], (n, v)
            if "field_waves" in vals:
                assert vals["field_waves"] in (4, 8)
                assert vals["field_wg_per_cu"] <= (2 if vals["field_waves"] == 8 else 3)


if __name__ == "__main__":
    pytest.main([__file__, "-q"])
