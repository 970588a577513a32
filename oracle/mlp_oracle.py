"""mlp_oracle.py - FP64 restatement of the proposal network (forward and
backward), TEST INFRASTRUCTURE ONLY (tests/).

Same parameter layout and semantics as cuacd_amd/csrc/mlp.hip: per
candidate plane, its part's 256 samples with their signed distance s to the
plane through a point MLP 4 -> 128 -> 128 -> 128 (ReLU), max-pooled over the
samples with s >= 0 and with s < 0 separately (0 for an empty side), head
[g+, g-, n, d] -> 128 -> 128 -> 1 (ReLU between), mean squared error. With
bf16=True the values the GPU stores in BF16 (inputs, weights, activations,
upstream gradients) are rounded to BF16 (round to nearest even) at the same
points, so the GPU's FP32-accumulated results differ from this FP64 oracle
only by accumulation rounding. No reference counterpart (the network is this
library's own); the backward pass is pinned by finite differences in FP64.
"""
import numpy as np

NS, HD, KIN, KH, NO = 256, 128, 16, 272, 16
SHAPES = [("W1", (KIN, HD)), ("b1", (HD,)), ("W2", (HD, HD)), ("b2", (HD,)), ("W3", (HD, HD)), ("b3", (HD,)),
          ("H1", (KH, HD)), ("c1", (HD,)), ("H2", (HD, HD)), ("c2", (HD,)), ("H3", (HD, NO)), ("c3", (NO,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, o = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = np.asarray(flat[o:o + n], np.float64).reshape(s)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, points, planes, part_of, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda x: np.asarray(x, np.float64))
    p = unpack(params)
    W = {k: (r(v) if k[0] in "WH" else v) for k, v in p.items()}
    pts = np.asarray(points, np.float32)
    pl = np.asarray(planes, np.float32)
    po = np.asarray(part_of)
    C = len(pl)
    P3 = pts[po]                                   # [C, NS, 3]
    # signed distances in FP32 as the GPU computes them (same operation order)
    s = ((pl[:, None, 0] * P3[:, :, 0] + pl[:, None, 1] * P3[:, :, 1]) + pl[:, None, 2] * P3[:, :, 2]) + pl[:, None, 3]
    X = np.zeros((C * NS, KIN))
    X[:, :3] = r(P3.reshape(-1, 3).astype(np.float64))
    X[:, 3] = r(s.reshape(-1).astype(np.float64))
    A1 = r(np.maximum(X @ W["W1"] + p["b1"], 0))
    A2 = r(np.maximum(A1 @ W["W2"] + p["b2"], 0))
    A3 = r(np.maximum(A2 @ W["W3"] + p["b3"], 0))
    A3c = A3.reshape(C, NS, HD)
    pos = (s >= 0)[:, :, None]
    neg = ~pos
    big = -1e30
    vp = np.where(pos, A3c, big)
    vn = np.where(neg, A3c, big)
    has_p = pos.any(axis=1)                        # [C, 1]
    has_n = neg.any(axis=1)
    gp = np.where(has_p, vp.max(axis=1), 0.0)
    gn = np.where(has_n, vn.max(axis=1), 0.0)
    ap = np.where(has_p, vp.argmax(axis=1), -1)    # first maximum, as the GPU's strict > scan
    an = np.where(has_n, vn.argmax(axis=1), -1)
    Hin = np.zeros((C, KH))
    Hin[:, :HD] = gp
    Hin[:, HD:2 * HD] = gn
    Hin[:, 2 * HD:2 * HD + 4] = pl.astype(np.float64)
    Hin = r(Hin)
    B1 = r(np.maximum(Hin @ W["H1"] + p["c1"], 0))
    B2 = r(np.maximum(B1 @ W["H2"] + p["c2"], 0))
    Y = B2 @ W["H3"] + p["c3"]
    cache = dict(X=X, A1=A1, A2=A2, A3=A3, ap=ap, an=an, Hin=Hin, B1=B1, B2=B2, Y=Y, W=W, C=C)
    return Y[:, 0], cache


def loss_and_grads(params, points, planes, part_of, targets, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda x: np.asarray(x, np.float64))
    y, c = forward(params, points, planes, part_of, emulate_bf16)
    W, C = c["W"], c["C"]
    t = np.asarray(targets, np.float64)
    e = y - t
    loss = float(np.mean(e * e))
    dY = np.zeros((C, NO))
    dY[:, 0] = 2.0 * e / C
    dY = r(dY)
    g = {}
    g["H3"] = c["B2"].T @ dY
    g["c3"] = dY.sum(0)
    dZ2 = r((dY @ W["H3"].T) * (c["B2"] > 0))
    g["H2"] = c["B1"].T @ dZ2
    g["c2"] = dZ2.sum(0)
    dZ1 = r((dZ2 @ W["H2"].T) * (c["B1"] > 0))
    g["H1"] = c["Hin"].T @ dZ1
    g["c1"] = dZ1.sum(0)
    dHin = dZ1 @ W["H1"].T
    dA3 = np.zeros((C, NS, HD))
    ci, ji = np.meshgrid(np.arange(C), np.arange(HD), indexing="ij")
    for arg, off in ((c["ap"], 0), (c["an"], HD)):
        ok = arg >= 0
        np.add.at(dA3, (ci[ok], arg[ok], ji[ok]), dHin[:, off:off + HD][ok])
    dA3 = dA3.reshape(C * NS, HD) * (c["A3"] > 0)
    dA3 = r(dA3)
    g["W3"] = c["A2"].T @ dA3
    g["b3"] = dA3.sum(0)
    dZ2e = r((dA3 @ W["W3"].T) * (c["A2"] > 0))
    g["W2"] = c["A1"].T @ dZ2e
    g["b2"] = dZ2e.sum(0)
    dZ1e = r((dZ2e @ W["W2"].T) * (c["A1"] > 0))
    g["W1"] = c["X"].T @ dZ1e
    g["b1"] = dZ1e.sum(0)
    return loss, pack(g), y
