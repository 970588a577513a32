"""Learned cut proposal for lookahead_decompose.

A point-set network (the library's BF16 MFMA kernels, mlp.hip) predicts the
exact cut score of each level-0 candidate plane of a part; with a network set
(Context.batch_lookahead_decompose(proposal=net, top_k=8)) the search
evaluates exactly only the top_k candidates of each cutting part, falling
back to the others when the kept ones yield no cut. Every accepted cut is
still the exact search's choice among the evaluated ones.

Training data are triples (part, plane, exact concavity) the exact pipeline
itself produces. record_dataset: plain lookahead runs on seeded SDF meshes
record, for every cutting part of every iteration, its samples and each
level-0 candidate's path cost - the exact value the search minimises when it
picks the cut. make_dataset: parts and their level-0 candidates scored by the
batched exact scorer (the larger max(0.3 rv, H) of the two halves). Failed
cuts score 1. Training runs forward, backward and Adam on the GPU (mean
squared error against each part's candidate ranks).
"""
import os
import time

import numpy as np

from . import _gpu

NS = 256            # surface samples per part
# weights trained by `python -m cuacd_amd.train_proposal` (shipped with the package)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "proposal_weights.npz")
FAIL_SCORE = 1.0    # target of a cut that fails or leaves a side empty


def n_params():
    return _gpu.mlp_n_params()


class ProposalNet:
    """The network on the device (FP32 master weights, BF16 compute copy)."""

    def __init__(self, ctx, params=None, seed=0):
        self.ctx = ctx
        p = None
        if params is not None:
            p = np.ascontiguousarray(params, np.float32).reshape(-1)
            if len(p) != n_params():
                raise ValueError(f"expected {n_params()} parameters, got {len(p)}")
        self.h = _gpu.mlp_create(p.ctypes.data if p is not None else 0, int(seed))

    def close(self):
        if self.h:
            _gpu.mlp_destroy(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(n_params(), np.float32)
        _gpu.mlp_get_params(self.h, out.ctypes.data)
        return out

    def save(self, path):
        np.savez(path, params=self.params())

    @classmethod
    def load_default(cls, ctx):
        return cls.load(ctx, DEFAULT_WEIGHTS)

    @classmethod
    def load(cls, ctx, path):
        with np.load(path, allow_pickle=False) as z:
            return cls(ctx, params=z["params"])

    @staticmethod
    def _inputs(points, planes, part_of):
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, NS, 3)
        pl = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
        po = np.ascontiguousarray(part_of, np.int32).reshape(-1)
        if len(po) != len(pl):
            raise ValueError("planes and part_of must have the same length")
        if len(po) and (po.min() < 0 or po.max() >= len(pts) or np.any(np.diff(po) < 0)):
            raise ValueError("part_of must be ascending part indices")
        return pts, pl, po

    def predict(self, points, planes, part_of):
        pts, pl, po = self._inputs(points, planes, part_of)
        out = np.empty(len(pl), np.float32)
        if len(pl):
            _gpu.mlp_forward(self.h, pts.ctypes.data, len(pts), pl.ctypes.data, po.ctypes.data, len(pl),
                             out.ctypes.data)
        return out

    def train_step(self, points, planes, part_of, targets, lr=1e-3, apply=True, return_grads=False):
        pts, pl, po = self._inputs(points, planes, part_of)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1)
        g = np.empty(n_params(), np.float32) if return_grads else None
        loss = _gpu.mlp_train_step(self.h, pts.ctypes.data, len(pts), pl.ctypes.data, po.ctypes.data, len(pl),
                                   t.ctypes.data, float(lr), int(apply), g.ctypes.data if g is not None else 0)
        return (loss, g) if return_grads else loss


def _pack(meshes):
    vs = [np.ascontiguousarray(v, np.float32).reshape(-1, 3) for v, _ in meshes]
    ts = [np.ascontiguousarray(t, np.int32).reshape(-1, 3) for _, t in meshes]
    voff = np.zeros(len(meshes) + 1, np.int32)
    toff = np.zeros(len(meshes) + 1, np.int32)
    voff[1:] = np.cumsum([len(v) for v in vs])
    toff[1:] = np.cumsum([len(t) for t in ts])
    return np.ascontiguousarray(np.concatenate(vs)), voff, np.ascontiguousarray(np.concatenate(ts)), toff


def sample_parts(ctx, meshes):
    """[P, 256, 3] area-stratified surface samples of each mesh in its box
    frame (center, half the largest extent) - the network's point input."""
    out = np.empty((len(meshes), NS, 3), np.float32)
    if meshes:
        v, vo, t, to = _pack(meshes)
        _gpu.mlp_sample_parts(v.ctypes.data, vo.ctypes.data, t.ctypes.data, to.ctypes.data, len(meshes),
                              out.ctypes.data)
    return out


def frame_planes(meshes, planes, part_of):
    """Planes (a, b, c, d) of world space in their part's box frame: unit
    normal, offset divided by the frame's scale (the search's own mapping)."""
    pl = np.asarray(planes, np.float32).reshape(-1, 4)
    out = np.empty_like(pl)
    for i, p in enumerate(np.asarray(part_of)):
        v = np.asarray(meshes[p][0], np.float32).reshape(-1, 3)
        lo, hi = v.min(0), v.max(0)
        c = np.float32(0.5) * (lo + hi)
        s = np.float32(max(float(np.max(np.float32(0.5) * (hi - lo))), 0.0)) or np.float32(1.0)
        n = pl[i, :3]
        ln = np.float32(np.sqrt(np.float32(n @ n)))
        il = np.float32(1.0) / ln if ln > 0 else np.float32(0.0)
        nn = n * il
        out[i, :3] = nn
        out[i, 3] = (pl[i, 3] * il + np.float32(nn @ c)) / s
    return out


def candidates(ctx, meshes, width=30, threshold=0.05, n_concave_edges=16, concave_eps=0.005,
               concave_threshold=3.49, cap=160):
    """The search's level-0 candidates of whole meshes: (planes [C, 4],
    part_of [C], k [C])."""
    n = len(meshes)
    v, vo, t, to = _pack(meshes)
    pl = np.zeros((n, cap, 4), np.float32)
    kk = np.zeros((n, cap), np.int32)
    cnt = np.zeros(n, np.int32)
    _gpu.lookahead_candidates(n, v.ctypes.data, vo.ctypes.data, t.ctypes.data, to.ctypes.data, int(width),
                              float(threshold), int(n_concave_edges), float(concave_eps), float(concave_threshold),
                              int(cap), pl.ctypes.data, kk.ctypes.data, cnt.ctypes.data)
    planes = np.concatenate([pl[m, :cnt[m]] for m in range(n)]) if n else np.zeros((0, 4), np.float32)
    part_of = np.concatenate([np.full(cnt[m], m, np.int32) for m in range(n)]) if n else np.zeros(0, np.int32)
    k = np.concatenate([kk[m, :cnt[m]] for m in range(n)]) if n else np.zeros(0, np.int32)
    return planes, part_of, k


def training_parts(ctx, meshes, iters=(0, 1, 2), chunk=128, **search):
    """Parts the search cuts: each mesh as it starts and the pieces of plain
    lookahead runs stopped after `iters` iterations (meshes in batches of
    `chunk`, which bounds the device heap one search uses)."""
    parts = []
    for k in iters:
        if k == 0:
            parts += [(np.asarray(v, np.float32), np.asarray(t, np.int32)) for v, t in meshes]
            continue
        for s in range(0, len(meshes), chunk):
            res = ctx.batch_lookahead_decompose(meshes[s:s + chunk], **dict(search, max_iters=k, merge_hulls=False))
            for ps in res:
                parts += [(pv, pt) for pv, pt, _, _ in ps if len(pt) >= 8]
    return parts


def record_dataset(ctx, meshes, chunk=128, **search):
    """(points, planes, part_of, targets) recorded by plain lookahead runs:
    every cutting part of every iteration with its level-0 candidates' path
    costs."""
    _gpu.lookahead_record(1)
    try:
        for s in range(0, len(meshes), chunk):
            ctx.batch_lookahead_decompose(meshes[s:s + chunk], **search)
        n_parts, n_cand = _gpu.lookahead_record_size()
        pts = np.empty((n_parts, NS, 3), np.float32)
        pl = np.empty((n_cand, 4), np.float32)
        po = np.empty(n_cand, np.int32)
        tg = np.empty(n_cand, np.float32)
        _gpu.lookahead_record_read(pts.ctypes.data, pl.ctypes.data, po.ctypes.data, tg.ctypes.data)
    finally:
        _gpu.lookahead_record(0)
    tg = np.where(tg >= 1e29, FAIL_SCORE, tg).astype(np.float32)
    return pts, pl, po, tg


def make_dataset(ctx, parts, threshold=0.05, width=30, n_concave_edges=16, chunk=256, log=None):
    """(points [P, 256, 3], planes [C, 4] in part frames, part_of [C], targets [C])."""
    pts, pls, pos, tgs = [], [], [], []
    base = 0
    for s in range(0, len(parts), chunk):
        ms = parts[s:s + chunk]
        planes, part_of, _ = candidates(ctx, ms, width=width, threshold=threshold,
                                        n_concave_edges=n_concave_edges)
        score = ctx.batch_score_cuts(ms, part_of, planes, threshold=threshold)
        score = np.where(score >= 1e29, FAIL_SCORE, score).astype(np.float32)
        pts.append(sample_parts(ctx, ms))
        pls.append(frame_planes(ms, planes, part_of))
        pos.append(part_of + base)
        tgs.append(score)
        base += len(ms)
        if log:
            log(f"[proposal] dataset {base}/{len(parts)} parts, {sum(len(x) for x in tgs)} triples")
    return (np.concatenate(pts), np.concatenate(pls), np.concatenate(pos).astype(np.int32),
            np.concatenate(tgs))


def _batches(part_of, n_parts, batch_parts, rng):
    order = rng.permutation(n_parts)
    starts = np.searchsorted(part_of, np.arange(n_parts + 1))
    for b in range(0, n_parts, batch_parts):
        sel = np.sort(order[b:b + batch_parts])
        idx = np.concatenate([np.arange(starts[p], starts[p + 1]) for p in sel])
        remap = np.empty(n_parts, np.int32)
        remap[sel] = np.arange(len(sel), dtype=np.int32)
        yield sel, idx, remap[part_of[idx]]


def rank_targets(part_of, scores):
    """Per part, each candidate's rank by exact score scaled to [0, 1] (0 =
    best; ties share the lower rank): what the proposal has to get right is
    the order inside a part, not the part's overall cost level."""
    t = np.zeros(len(scores), np.float32)
    starts = np.searchsorted(part_of, np.arange(part_of.max() + 2)) if len(part_of) else [0]
    for a, b in zip(starts[:-1], starts[1:]):
        if b - a <= 1:
            continue
        s = scores[a:b]
        r = np.searchsorted(np.sort(s, kind="stable"), s, side="left")
        t[a:b] = r / float(b - a - 1)
    return t


def train(ctx, data, epochs=30, batch_parts=64, lr=1e-3, seed=0, net=None, log=None, target="rank"):
    """Adam on the mean squared error against per-part rank targets (or the
    raw scores with target="score"); returns (net, per-epoch mean losses)."""
    pts, pl, po, tg = data
    if target == "rank":
        tg = rank_targets(po, tg)
    rng = np.random.default_rng(seed)
    net = net or ProposalNet(ctx, seed=seed)
    hist = []
    for ep in range(epochs):
        t0 = time.perf_counter()
        losses = []
        for sel, idx, po_b in _batches(po, len(pts), batch_parts, rng):
            if len(idx) == 0:
                continue
            losses.append(net.train_step(pts[sel], pl[idx], po_b, tg[idx], lr=lr))
        hist.append(float(np.mean(losses)))
        if log:
            log(f"[proposal] epoch {ep}: loss {hist[-1]:.5f} ({time.perf_counter() - t0:.2f} s)")
    return net, hist


def ranking_quality(net, data, top_k=8):
    """Fraction of parts whose best exact candidate is among the predicted top_k,
    and the mean regret (best kept exact score minus best exact score)."""
    pts, pl, po, tg = data
    pred = net.predict(pts, pl, po)
    starts = np.searchsorted(po, np.arange(len(pts) + 1))
    hit, regret, n = 0, 0.0, 0
    for p in range(len(pts)):
        a, b = starts[p], starts[p + 1]
        if b - a <= top_k:
            continue
        keep = a + np.argsort(pred[a:b], kind="stable")[:top_k]
        best = tg[a:b].min()
        hit += int(tg[keep].min() <= best)
        regret += float(tg[keep].min() - best)
        n += 1
    return (hit / max(n, 1), regret / max(n, 1), n)
