"""CPU tests of the autotuner's search and of the launch-shape knobs' host
side (set/get with clamping; no GPU needed)."""
import numpy as np

from cuacd_amd import autotune
import cuacd_amd


class FakeCtx:
    def __init__(self):
        self.t = {"hull_wide_min": 256, "hull_vcap_eighths": 3, "hull_half_cap": 1, "kdop_min_points": 512,
                  "hull_chunk_above": 16384, "hull_multi": 1, "hull_ldsv8_vcap": 1024,
                  "hull_ldsb8_vcap": 1024, "cut_lds_verts": 98304, "hull_lpt": 0,
                  "hull_b4_split": 0, "hull_order": 1, "hull_w8": 5, "hull_multi_hbm": 1}

    def get_tuning(self, name=None):
        return dict(self.t) if name is None else self.t[name]

    def set_tuning(self, name, value):
        self.t[name] = value


def test_tuner_picks_fastest_setting_with_same_output(monkeypatch):
    ctx = FakeCtx()
    # simulated cost: vcap 4 is fastest but changes the output (must be
    # rejected); kdop 512 is faster and output-preserving
    cost = {("hull_vcap_eighths", 4): 0.1, ("kdop_min_points", 512): 0.5}
    clock = [0.0]

    def fake_time():
        return clock[0]

    def run():
        c = 1.0
        for k, v in ctx.t.items():
            c = min(c, cost.get((k, v), 1.0))
        clock[0] += c
        return np.array([ctx.t["hull_vcap_eighths"] == 4], bool)

    monkeypatch.setattr(autotune.time, "perf_counter", fake_time)
    best, trials = autotune.tune(ctx, run)
    assert best["kdop_min_points"] == 512 and best["hull_vcap_eighths"] == 3
    assert ctx.t == best
    rejected = [s for s, _, ok in trials if not ok]
    assert rejected and all(s["hull_vcap_eighths"] == 4 for s in rejected)


def test_tuning_knobs_round_trip_and_clamp():
    g = cuacd_amd._gpu
    keys = cuacd_amd.TUNING_KEYS
    old = {k: g.get_tuning(v) for k, v in keys.items()}
    try:
        g.set_tuning(keys["hull_vcap_eighths"], 99)
        assert g.get_tuning(keys["hull_vcap_eighths"]) == 8
        g.set_tuning(keys["kdop_min_points"], 1)
        assert g.get_tuning(keys["kdop_min_points"]) == 4
        g.set_tuning(keys["hull_half_cap"], 5)
        assert g.get_tuning(keys["hull_half_cap"]) == 1
        g.set_tuning(keys["hull_wide_min"], 512)
        assert g.get_tuning(keys["hull_wide_min"]) == 512
        g.set_tuning(keys["hull_chunk_above"], 100000)
        assert g.get_tuning(keys["hull_chunk_above"]) == 16384
        g.set_tuning(keys["hull_chunk_above"], 10)
        assert g.get_tuning(keys["hull_chunk_above"]) == 1024
    finally:
        for k, v in old.items():
            g.set_tuning(keys[k], v)


def test_cache_round_trip(tmp_path):
    class C(FakeCtx):
        def device_info(self):
            return ("gfx950", 256, 1 << 38)

    ctx = C()
    p = str(tmp_path / "t.json")
    assert not autotune.load(ctx, "w", path=p)
    autotune.save(ctx, "w", {"kdop_min_points": 512}, path=p)
    assert autotune.load(ctx, "w", path=p) and ctx.t["kdop_min_points"] == 512
