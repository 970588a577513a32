"""Autotuner for the hull tiers' launch shapes.

The knobs (Context.set_tuning: the four-wave threshold, the first-pass vertex
capacity, the reduced-capacity first passes, the k-DOP prefilter threshold)
choose kernels and LDS footprints, never results: every tier computes the
same canonical hull. The tuner therefore times a caller-supplied workload
under each candidate setting, checks that its output is bit-identical to the
output under the starting setting (a setting that changes it is rejected and
reported), and keeps the fastest. Search: coordinate descent over the knobs in
order, starting from the current values, `repeats` timed runs per setting
(after one untimed run of the starting setting). Results can be cached per
device and workload key in a JSON file.
"""
import json
import os
import time

import numpy as np

DEFAULT_SPACE = {
    "hull_wide_min": [128, 256, 512],
    "hull_vcap_eighths": [2, 3, 4],
    "hull_half_cap": [0, 1],
    "kdop_min_points": [128, 256, 512],
    "hull_multi": [0, 1],
    "hull_ldsv8_vcap": [448, 1024],
    "hull_ldsb8_vcap": [0, 1024],
    "hull_lpt": [0, 1],
    "hull_order": [0, 1],
    "hull_w8": [0, 1, 5],
    "hull_multi_hbm": [0, 1],
}

CACHE_PATH = os.path.join(os.path.expanduser("~"), ".cache", "cuacd_amd", "tuning.json")


def _same(a, b):
    if isinstance(a, (tuple, list)):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    return np.array_equal(np.asarray(a), np.asarray(b))


def tune(ctx, run, space=None, repeats=1, log=None):
    """Tune the knobs for `run()` (a callable doing the workload on ctx and
    returning its output). Applies and returns (best settings, trial records);
    each record is (settings, seconds, accepted)."""
    space = DEFAULT_SPACE if space is None else space
    best = ctx.get_tuning()
    for k, v in best.items():
        ctx.set_tuning(k, v)
    ref = run()

    def timed(settings):
        for k, v in settings.items():
            ctx.set_tuning(k, v)
        t = []
        out = None
        for _ in range(max(1, repeats)):
            t0 = time.perf_counter()
            out = run()
            t.append(time.perf_counter() - t0)
        return min(t), out

    best_t, _ = timed(best)
    trials = [(dict(best), best_t, True)]
    for knob, values in space.items():
        for v in values:
            if v == best[knob]:
                continue
            cand = dict(best)
            cand[knob] = v
            t, out = timed(cand)
            ok = _same(out, ref)
            trials.append((cand, t, ok))
            if log:
                log(f"[autotune] {cand} {t * 1e3:.1f} ms{'' if ok else ' (output differs: rejected)'}")
            if ok and t < best_t:
                best, best_t = cand, t
    for k, v in best.items():
        ctx.set_tuning(k, v)
    return best, trials


def device_key(ctx):
    name, n_cu, hbm = ctx.device_info()
    return f"{name}|{n_cu}|{hbm}"


def load(ctx, workload, path=CACHE_PATH):
    """Apply cached settings for (device, workload); True when found."""
    try:
        with open(path) as f:
            cache = json.load(f)
    except (OSError, ValueError):
        return False
    s = cache.get(device_key(ctx), {}).get(workload)
    if not s:
        return False
    for k, v in s.items():
        ctx.set_tuning(k, v)
    return True


def save(ctx, workload, settings, path=CACHE_PATH):
    try:
        with open(path) as f:
            cache = json.load(f)
    except (OSError, ValueError):
        cache = {}
    cache.setdefault(device_key(ctx), {})[workload] = dict(settings)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
