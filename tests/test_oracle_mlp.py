"""CPU pins of the proposal network's FP64 oracle (oracle/mlp_oracle.py):
its backward pass against central finite differences of its own loss in
exact FP64 (no BF16 rounding), on a small random instance; a planted
mistake (a transposed head weight in the backward) must fail the check."""
import numpy as np

import mlp_oracle as MO


def _instance(seed=0, P=2, C=7):
    rng = np.random.default_rng(seed)
    p = MO.unpack(np.zeros(MO.N_PARAMS))
    for name, s in MO.SHAPES:
        p[name] = rng.normal(0, 0.3, s)
    p["W1"][4:] = 0
    p["H1"][260:] = 0
    p["H3"][:, 1:] = 0
    p["c3"][1:] = 0
    params = MO.pack(p)
    pts = rng.uniform(-1, 1, (P, MO.NS, 3))
    planes = rng.normal(size=(C, 4))
    part_of = np.sort(rng.integers(0, P, C))
    t = rng.uniform(0, 0.5, C)
    return params, pts, planes, part_of, t


def _fd_check(grad_fn, params, args, idx, h=1e-6):
    _, g, _ = grad_fn(params, *args, emulate_bf16=False)
    worst = 0.0
    for i in idx:
        pp = params.copy(); pp[i] += h
        pm = params.copy(); pm[i] -= h
        lp = grad_fn(pp, *args, emulate_bf16=False)[0]
        lm = grad_fn(pm, *args, emulate_bf16=False)[0]
        fd = (lp - lm) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(1e-8, abs(fd) + abs(g[i])))
    return worst


def _sample_idx(params, rng, k=6):
    o, idx = 0, []
    for name, s in MO.SHAPES:
        n = int(np.prod(s))
        blk = MO.unpack(np.arange(MO.N_PARAMS, dtype=np.float64))[name].reshape(-1).astype(int)
        used = blk[np.abs(params[blk]) > 0] if name[0] in "WH" else blk
        idx += list(rng.choice(used, min(k, len(used)), replace=False))
        o += n
    return idx


def test_backward_matches_finite_differences():
    params, pts, planes, part_of, t = _instance()
    rng = np.random.default_rng(1)
    idx = _sample_idx(params, rng)
    worst = _fd_check(MO.loss_and_grads, params, (pts, planes, part_of, t), idx)
    assert worst < 1e-5, worst


def test_planted_transpose_is_caught():
    params, pts, planes, part_of, t = _instance(seed=2)

    def bad(params_, pts_, planes_, part_of_, t_, emulate_bf16=False):
        loss, g, y = MO.loss_and_grads(params_, pts_, planes_, part_of_, t_, emulate_bf16)
        d = MO.unpack(g)
        d["H2"] = d["H2"].T.copy()  # the mistake
        return loss, MO.pack(d), y

    rng = np.random.default_rng(3)
    h2 = MO.unpack(np.arange(MO.N_PARAMS, dtype=np.float64))["H2"].reshape(-1).astype(int)
    worst = _fd_check(bad, params, (pts, planes, part_of, t), list(rng.choice(h2, 8, replace=False)))
    assert worst > 1e-2


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], np.float64)
    r = MO.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
