"""lookahead_oracle.py - serial restatement of lookahead_decompose.

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU baseline leg).

Follows the reference search (csrc/lookahead.c:144-1124; cuda/la_lifecycle.cu,
cuda/la_expand.cu, cuda/la_refine.cu, cuda/la_concave.cu, cuda/la_axis_cut.cuh,
cuda/la_common.cuh) with the GPU implementation's canonical choices, so that
the two produce the same decomposition on general-position inputs:
  * volumes in FP64, costs rounded to float32 (cost_rv);
  * the plane cut, the leaf scoring and the concave-edge records come from
    the C oracle (cut_oracle.c, lookahead_oracle.c);
  * hull volumes from the exact quantized hull of the full point set; sets
    above 1024 points first get the rough hull of their 80 k-DOP extreme
    points, whose volume drives the branch-and-bound pruning;
  * Hausdorff distances with the reference sampler and exact FP64 distances;
  * new parts are appended in cutting order.
Slow (pure Python control, serial C pieces): meant for small meshes.
"""
import math

import numpy as np

import oracle as O

LA_MAX_PARTS = 8
LA_MAX_LEVELS = 4
PI_F = np.float32(3.14159265358979)


def cost_rv(hv, mv):
    g = hv - mv
    g = g if g > 0.0 else 0.0
    return np.float32(0.3 * np.cbrt(3.0 / (4.0 * math.pi) * g))


def mesh_volume_abs(v, t):
    """FP64 |signed volume| anchored at the first vertex (the GPU's pc_volume)."""
    if len(t) == 0:
        return 0.0
    v = np.asarray(v, np.float32)
    a0 = v[0].astype(np.float64)
    p = v[np.asarray(t)].astype(np.float64) - a0
    x0, x1, x2 = p[:, 0], p[:, 1], p[:, 2]
    det = (x0[:, 0] * (x1[:, 1] * x2[:, 2] - x1[:, 2] * x2[:, 1]) + x0[:, 1] * (x1[:, 2] * x2[:, 0] - x1[:, 0] * x2[:, 2]) +
           x0[:, 2] * (x1[:, 0] * x2[:, 1] - x1[:, 1] * x2[:, 0]))
    return abs(float(det.sum())) / 6.0


def signed_volume_first(v, t):
    v = np.asarray(v, np.float32)
    a0 = v[0].astype(np.float64)
    p = v[np.asarray(t)].astype(np.float64) - a0
    x0, x1, x2 = p[:, 0], p[:, 1], p[:, 2]
    det = (x0[:, 0] * (x1[:, 1] * x2[:, 2] - x1[:, 2] * x2[:, 1]) + x0[:, 1] * (x1[:, 2] * x2[:, 0] - x1[:, 0] * x2[:, 2]) +
           x0[:, 2] * (x1[:, 0] * x2[:, 1] - x1[:, 1] * x2[:, 0]))
    return float(det.sum()) / 6.0


# ---- k-DOP axes: unit normals of the level-1 icosphere's faces, one per antipodal pair
def kdop_axes():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    v = [[c / math.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) for c in p] for p in v]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]

    def mid(a, b):
        m = [v[a][i] + v[b][i] for i in range(3)]
        ln = math.sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])
        m = [c / ln for c in m]
        for i, p in enumerate(v):
            if abs(p[0] - m[0]) + abs(p[1] - m[1]) + abs(p[2] - m[2]) < 1e-12:
                return i
        v.append(m)
        return len(v) - 1

    normals = []
    for a, b, c in F:
        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        for s in ((a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)):
            p, q, r = v[s[0]], v[s[1]], v[s[2]]
            ux, uy, uz = q[0] - p[0], q[1] - p[1], q[2] - p[2]
            wx, wy, wz = r[0] - p[0], r[1] - p[1], r[2] - p[2]
            n = [uy * wz - uz * wy, uz * wx - ux * wz, ux * wy - uy * wx]
            ln = math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
            n = [c2 / ln for c2 in n]
            if not any(abs(o[0] + n[0]) + abs(o[1] + n[1]) + abs(o[2] + n[2]) < 1e-9 for o in normals):
                normals.append(n)
    assert len(normals) == 40
    return np.array(normals, np.float32)


_AXES = None


def hull_volume(pts, prune_gap=1e30, mesh_vol=0.0):
    """(volume, pruned) of the exact hull of pts, with the k-DOP rough-hull
    pruning for sets above 1024 points (reference cuda/kdop_hull.cuh:49-398)."""
    global _AXES
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    n = len(pts)
    if n < 4:
        return 0.0, False
    if n > 1024:
        if _AXES is None:
            _AXES = kdop_axes()
        A = _AXES
        d = (A[:, 0:1] * pts[:, 0][None, :] + A[:, 1:2] * pts[:, 1][None, :]) + A[:, 2:3] * pts[:, 2][None, :]
        imax = np.argmax(d, axis=1)
        imin = np.argmin(d, axis=1)
        ep = np.empty((80, 3), np.float32)
        ep[0::2] = pts[imax]
        ep[1::2] = pts[imin]
        flags, faces = O.hull(O.quantize(ep))
        if len(faces) and prune_gap < 1e29:
            rough = abs(O.hull_volume(ep, faces))
            if max(rough - mesh_vol, 0.0) > prune_gap:
                return rough, True
    flags, faces = O.hull(O.quantize(pts))
    if len(faces) == 0:
        return 0.0, False
    return abs(O.hull_volume(pts, faces)), False


def hull_mesh(pts):
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    if len(pts) < 4:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), 0.0
    flags, faces = O.hull(O.quantize(pts))
    if len(faces) == 0:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), 0.0
    idx = np.nonzero(flags)[0]
    remap = -np.ones(len(pts), np.int64)
    remap[idx] = np.arange(len(idx))
    return pts[idx].copy(), remap[faces].astype(np.int32), abs(O.hull_volume(pts, faces))


def hausdorff(hv, ht, mv, mt):
    if len(ht) == 0 or len(mt) == 0:
        return np.float32(0.0)
    sa = O.reference_samples(hv, ht)
    sb = O.reference_samples(mv, mt)
    d = max(O.directed_hausdorff(sa, mv, mt), O.directed_hausdorff(sb, hv, ht))
    return np.float32(d)


def components(v, t):
    """Pieces by union-find over triangle vertices, numbered by lowest vertex;
    vertices in index order, triangles in input order; inner shells (negative
    signed volume) dropped when mixed with outer ones (cuda/postprocess.cuh)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    nv = len(v)
    used = np.zeros(nv, bool)
    used[t.reshape(-1)] = True
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]]])
    g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(nv, nv))
    _, lab = connected_components(g, directed=False)
    roots = {}
    order = []
    for i in range(nv):
        if used[i] and lab[i] not in roots:
            roots[lab[i]] = len(order)
            order.append(lab[i])
    if len(order) <= 1:
        return None
    cid = np.array([roots.get(l, -1) if used[i] else -1 for i, l in enumerate(lab)])
    cid = np.minimum(cid, 255)
    out = []
    for c in range(min(len(order), 256)):
        vi = np.nonzero(cid == c)[0]
        loc = -np.ones(nv, np.int64)
        loc[vi] = np.arange(len(vi))
        tm = cid[t[:, 0]] == c
        cv = v[vi].copy()
        ct = loc[t[tm]].astype(np.int32)
        out.append((cv, ct, signed_volume_first(cv, ct)))
    inner = [o[2] < 0 for o in out]
    if any(inner) and not all(inner):
        out = [o for o in out if o[2] >= 0]
    return out


def axis_cut(width, k):
    count, rem = width // 3, width % 3
    xc, yc = count + (rem > 0), count + (rem > 1)
    if k < xc:
        return 0, k, xc
    if k < xc + yc:
        return 1, k - xc, yc
    return 2, k - xc - yc, count


class Part:
    __slots__ = ("v", "t", "hv", "ht", "mvol", "hvol", "haus", "cc", "lo", "hi")

    def __init__(self, v, t, mvol, hvol=0.0, hv=None, ht=None, cc=0):
        self.v, self.t = v, t
        self.mvol, self.hvol = mvol, hvol
        self.hv, self.ht = hv, ht
        self.haus = np.float32(0.0)
        self.cc = cc
        self.lo, self.hi = v.min(0), v.max(0)

    def cost(self):
        return max(cost_rv(self.hvol, self.mvol), self.haus)


def concave_planes(p, n_concave, eps, thr):
    import ctypes as C
    L = O.lib()
    if not getattr(L, "_ce", False):
        L.orc_concave_edges.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_int,
                                        C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_concave_edges.restype = C.c_int
        L._ce = True
    x1, x2 = thr - math.pi, math.pi - thr
    if x1 < 0:
        pm, cpos = 1, 0.0
    elif x1 >= math.pi:
        pm, cpos = 0, 0.0
    else:
        pm, cpos = 2, float(np.float32(math.cos(x1)))
    if x2 <= 0:
        nm, cneg = 0, 0.0
    elif x2 > math.pi:
        nm, cneg = 1, 0.0
    else:
        nm, cneg = 2, float(np.float32(math.cos(x2)))
    v = np.ascontiguousarray(p.v, np.float32)
    t = np.ascontiguousarray(p.t, np.int32)
    cap = 3 * len(t) // 2 + 16
    ri = np.zeros((cap, 4), np.int32)
    rf = np.zeros((cap, 9), np.float32)
    cnt = L.orc_concave_edges(O._p(v), len(v), O._p(t), len(t), cpos, cneg, pm, nm, O._p(ri), O._p(rf), cap)
    cnt = min(cnt, cap)
    order = sorted(range(cnt), key=lambda e: (ri[e, 0], ri[e, 1], ri[e, 2]))
    res = [order[j] for j in reservoir(len(order), n_concave)]
    planes = []
    f = np.float32
    eps = f(eps)
    for e in res:
        n0, n1, m = rf[e, 0:3], rf[e, 3:6], rf[e, 6:9]
        planes.append([n1[0], n1[1], n1[2], -(((n1[0] * m[0] + n1[1] * m[1]) + n1[2] * m[2]) + eps)])
        planes.append([n0[0], n0[1], n0[2], -(((n0[0] * m[0] + n0[1] * m[1]) + n0[2] * m[2]) + eps)])
        bx, by, bz = f(n0[0] + n1[0]), f(n0[1] + n1[1]), f(n0[2] + n1[2])
        bl = f(np.sqrt(f(f(bx * bx + by * by) + bz * bz)))
        if bl >= f(1e-4):
            bx, by, bz = f(bx / bl), f(by / bl), f(bz / bl)
            dd = f(f(bx * m[0] + by * m[1]) + bz * m[2])
            planes.append([bx, by, bz, -(dd + eps)])
            planes.append([bx, by, bz, -(dd - eps)])
    return [np.array(pl, np.float32) for pl in planes]


def cut_cost(v, t, plane, tau):
    """Exact cost of one candidate cut, serially: the full plane cut, the exact
    hull of each closed half, cost_rv of each half (mesh volume clamped to the
    hull volume, cuda/la_lifecycle.cu:292-296), its Hausdorff distance to its
    hull when cost_rv < tau (cuda/la_refine.cu:19-45), the half's cost
    max(cost_rv, H) (cuda/la_common.cuh:21-34) and the cut's score as the
    larger of the two. Returns (score, [(mvol, hvol, crv, H or -1)] * 2);
    score 1e30 when the cut fails or leaves a side empty."""
    st, pv, pt, nv, nt = O.plane_cut(v, t, plane)
    if st != 0 or len(pv) == 0 or len(nv) == 0:
        return np.float32(1e30), None
    halves = []
    for hv_, ht_ in ((pv, pt), (nv, nt)):
        mv = mesh_volume_abs(hv_, ht_)
        hmv, hmt, hvol = hull_mesh(hv_)
        if mv > hvol:
            mv = hvol
        crv = cost_rv(hvol, mv)
        h = np.float32(-1.0)
        if crv < np.float32(tau) and len(hmt):
            h = hausdorff(hmv, hmt, hv_, ht_)
        halves.append((mv, hvol, crv, h))
    cost = [max(c, h if h > 0 else np.float32(0.0)) for _, _, c, h in halves]
    return np.float32(max(cost)), halves


def score_cut(v, t, plane):
    import ctypes as C
    L = O.lib()
    if not getattr(L, "_sc", False):
        L.orc_score_cut.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int] + [C.c_float] * 4 + \
            [C.c_void_p] * 4
        L._sc = True
    v = np.ascontiguousarray(v, np.float32)
    t = np.ascontiguousarray(t, np.int32)
    vol = np.zeros(2, np.float64)
    cnt = np.zeros(4, np.int32)
    pp = np.zeros((len(v) + 3 * len(t) + 1, 3), np.float32)
    npp = np.zeros((len(v) + 3 * len(t) + 1, 3), np.float32)
    L.orc_score_cut(O._p(v), len(v), O._p(t), len(t), *[float(x) for x in plane], O._p(vol), O._p(cnt), O._p(pp),
                    O._p(npp))
    return vol[0], vol[1], int(cnt[0]), int(cnt[1]), pp[:cnt[2]].copy(), npp[:cnt[3]].copy()


class Item:
    __slots__ = ("src", "cut0", "parts", "lc", "pruned", "store")

    def __init__(self, src, cut0, parts, lc, store=None):
        self.src, self.cut0, self.parts, self.lc = src, cut0, parts, lc
        self.pruned = False
        self.store = store  # (pos Part, neg Part) of a level-0 cut


def _sort_parts(parts):
    # stable insertion sort by rv cost (cuda/la_refine.cu:376-387)
    out = list(parts)
    for i in range(1, len(out)):
        key = out[i]
        kc = cost_rv(key[1], key[0])
        j = i - 1
        while j >= 0 and cost_rv(out[j][1], out[j][0]) > kc:
            out[j + 1] = out[j]
            j -= 1
        out[j + 1] = key
    return out


def level_sum(lc):
    """Sum of an item's level costs in level order, in float32."""
    ps = np.float32(0)
    for x in lc:
        ps = np.float32(ps + x)
    return ps


def path_cost(lc):
    """Path cost of a leaf: the mean of its level costs, summed in level order
    and divided in float32 (cuda/la_refine.cu:450-457)."""
    return np.float32(level_sum(lc) / np.float32(len(lc)))


def item_ub(lc, total):
    """Branch-and-bound upper bound of one item: its level costs so far plus
    its last level cost for every level still to come, since level costs do
    not increase along a path (cuda/la_expand.cu:434-457)."""
    return np.float32(level_sum(lc) + np.float32(lc[-1] * np.float32(total - len(lc))))


def best_ub(lcs, total):
    """Smallest item_ub over items with at least one level (1e30 if none)."""
    b = np.float32(1e30)
    for lc in lcs:
        if len(lc) > 0:
            b = min(b, item_ub(lc, total))
    return b


def prune_gap(lc, bub, total):
    """Rough-hull volume gap above which an item's new parts are pruned
    (cuda/la_expand.cu:380-399): the remaining cost budget bub - sum(lc) as
    an rv cost, turned back into the volume of a sphere of radius
    budget / 0.3. 1e30 disables pruning (no bound, or an item with no level
    or already at the last level), -1 always prunes (no budget left)."""
    nl = len(lc)
    if not (bub < np.float32(1e29) and 0 < nl < total):
        return 1e30
    rv_thr = np.float32(bub - level_sum(lc))
    if rv_thr <= 0:
        return -1.0
    r = np.float32(rv_thr / np.float32(0.3))
    return float(np.float32(np.float32(np.float32(r * r) * r) *
                            np.float32(np.float32(np.float32(4.0) * PI_F) / np.float32(3.0))))


def select_cutting(costs, tau, max_n_cutting):
    """The parts cut in one iteration (cuda/la_lifecycle.cu:90-144): a stable
    ascending sort of the parts by full cost, then the parts at or above tau,
    which form a suffix of that order; when more than max_n_cutting are above,
    the last (costliest) max_n_cutting. Returns (order, cutting): order is the
    sorted permutation of the input, cutting the positions in that order."""
    order = sorted(range(len(costs)), key=lambda i: costs[i])
    c = [costs[i] for i in order]
    n = len(c)
    first = next((i for i in range(n) if c[i] >= tau), n)
    start = n - max_n_cutting if n - first > max_n_cutting else first
    return order, [i for i in range(start, n) if c[i] >= tau][:max_n_cutting]


def per_cut_best(tw0, paths):
    """Smallest path cost per first cut index (1e30 where no leaf is left)."""
    best = [np.float32(1e30)] * tw0
    for cut0, pc in paths:
        if 0 <= cut0 < tw0 and pc < best[cut0]:
            best[cut0] = pc
    return best


def evaluate(tw0, paths, level0):
    """Best first cut of one cutting part (cuda/la_refine.cu:418-527).
    paths: (first cut index, path cost) of every leaf that was not pruned
    (full leaves, scored last-level children, extra leaves); each first cut
    keeps its smallest path cost. When no path is left (every subtree pruned
    or cut-less), the fallback is the greedy one: each valid level-0 item's
    own level-0 cost, level0 = {first cut index: level-0 cost}. Returns
    (best cost, first cut index), the lowest index among equal costs; index
    None when nothing is below 1e30."""
    best = per_cut_best(tw0, paths)
    if not any(b < np.float32(1e30) for b in best):
        for k, c0 in level0.items():
            best[k] = c0
    bc, bi = np.float32(1e30), None
    for k in range(tw0):
        if best[k] < bc:
            bc, bi = best[k], k
    return bc, bi


def lcg_next(s):
    """The reservoir's linear congruential step (Numerical Recipes constants;
    cuda/la_concave.cu lcg_next)."""
    return (s * 1664525 + 1013904223) & 0xFFFFFFFF


def reservoir(n_edges, k):
    """Reservoir sample of k of n_edges concave edges in their scan order
    (cuda/la_concave.cu:276-301): the first k fill the slots; afterwards the
    seen-th edge draws r from the LCG (seeded with 1) and replaces slot
    (next draw) mod k when r < (uint32)(float(k) / float(seen) * 4294967295f).
    Returns the scan indices held by the slots, in slot order."""
    res, rng = [], 1
    for seen in range(1, n_edges + 1):
        e = seen - 1
        if len(res) < k:
            res.append(e)
            continue
        rng = lcg_next(rng)
        thr = int(np.float32(np.float32(k) / np.float32(seen)) * np.float32(4294967295.0))
        if rng < thr:
            rng = lcg_next(rng)
            res[rng % k] = e
    return res


def quick_levels(cur, extra, quick_depth, total, med):
    """Quick levels (csrc/lookahead.c:656-727; cuda/la_refine.cu:52-351): one
    child per item, its worst part cut at the box midpoint along the axis with
    the most balanced halves, cost max(v+, v-) / (v+ + v- + 1e-10) in float32,
    the first axis on ties. Axes spanning <= 2 med + 1e-6 are skipped; an item
    with every axis skipped becomes an extra leaf (level costs padded with
    zeros), one with no valid cut is dropped, and a level with no child at
    all leaves the items as they are. The last quick level prunes with the
    bound over its input items. Returns the items that end the search."""
    ext = np.float32(np.float32(2.0) * med + np.float32(1e-6))
    for q in range(quick_depth):
        bub = best_ub([itm.lc for itm in cur], total) if q == quick_depth - 1 else np.float32(1e30)
        nxt = []
        for itm in cur:
            wm, wh, wv, wt, wlo, whi = itm.parts[-1]
            assert 0 < len(wv) <= 1000000 and 0 < len(wt) <= 1000000
            span = [np.float32(whi[j] - wlo[j]) for j in range(3)]
            if all(sp <= ext for sp in span):
                lc = list(itm.lc) + [np.float32(0)] * (min(total, LA_MAX_LEVELS) - len(itm.lc))
                extra.append(Item(itm.src, itm.cut0, itm.parts, lc))
                continue
            best, bc = None, np.float32(1e30)
            for ax in range(3):
                if span[ax] <= ext:
                    continue
                pl = np.zeros(4, np.float32)
                pl[ax] = 1.0
                pl[3] = -np.float32(np.float32(wlo[ax] + whi[ax]) * np.float32(0.5))
                st, pv, pt, nv, nt = O.plane_cut(wv, wt, pl)
                if st != 0 or len(pv) == 0 or len(nv) == 0:
                    continue
                pm, nmv = mesh_volume_abs(pv, pt), mesh_volume_abs(nv, nt)
                fp, fn = np.float32(pm), np.float32(nmv)
                cost = np.float32(max(fp, fn) / np.float32(np.float32(fp + fn) + np.float32(1e-10)))
                if cost < bc:
                    bc, best = cost, (pm, pv, pt, nmv, nv, nt)
            if best is None:
                continue
            assert len(itm.parts) + 1 <= LA_MAX_PARTS
            pm, pv, pt, nmv, nv, nt = best
            gap = prune_gap(itm.lc, bub, total)
            ph, prp = hull_volume(pv, gap, pm)
            nh, prn = hull_volume(nv, gap, nmv)
            ch = _sort_parts(list(itm.parts[:-1]) + [(pm, ph, pv, pt, pv.min(0), pv.max(0)),
                                                     (nmv, nh, nv, nt, nv.min(0), nv.max(0))])
            lc = list(itm.lc)
            if len(lc) < LA_MAX_LEVELS:
                lc.append(cost_rv(ch[-1][1], ch[-1][0]))
            c_item = Item(itm.src, itm.cut0, ch, lc)
            c_item.pruned = bool(prp or prn)
            nxt.append(c_item)
        if not nxt:
            break
        cur = nxt
    return cur


MERGE_INVALID, MERGE_RV_REJECT, MERGE_BBOX_REJECT = np.float32(1e20), np.float32(1e10), np.float32(1e9)


def shared_plane(av, at, bv, tol=np.float32(1e-3)):
    """First face of hull A (in face order) whose outward plane has all of A's
    hull vertices within tol on its inner side and all of B's hull vertices
    within tol of it (cuda/merge.cuh:111-165), as (nx, ny, nz, d) in float32;
    None if there is none."""
    f = np.float32
    av = np.asarray(av, np.float32)
    bv = np.asarray(bv, np.float32)
    for tri in np.asarray(at, np.int64):
        a, b, c = av[tri[0]], av[tri[1]], av[tri[2]]
        e1 = [f(b[j] - a[j]) for j in range(3)]
        e2 = [f(c[j] - a[j]) for j in range(3)]
        nx = f(f(e1[1] * e2[2]) - f(e1[2] * e2[1]))
        ny = f(f(e1[2] * e2[0]) - f(e1[0] * e2[2]))
        nz = f(f(e1[0] * e2[1]) - f(e1[1] * e2[0]))
        ln = f(np.sqrt(f(f(f(nx * nx) + f(ny * ny)) + f(nz * nz))))
        if ln < f(1e-20):
            continue
        inv = f(f(1.0) / ln)
        nx, ny, nz = f(nx * inv), f(ny * inv), f(nz * inv)
        d = f(f(f(nx * a[0]) + f(ny * a[1])) + f(nz * a[2]))
        sa = ((nx * av[:, 0] + ny * av[:, 1]) + nz * av[:, 2]) - d
        if (sa > tol).any():
            continue
        sb = ((nx * bv[:, 0] + ny * bv[:, 1]) + nz * bv[:, 2]) - d
        if (np.abs(sb) > tol).any():
            continue
        return nx, ny, nz, d
    return None


def merge_match(cost, tau):
    """Greedy matching of the reference (cuda/postprocess_merge.cu:404-480):
    256 threads each take the first minimum of their stride of the pair costs
    (below tau), a tree reduction keeps the left operand unless the right one
    is strictly smaller, the winning pair is recorded and every pair sharing a
    part with it is invalidated; until no pair is below tau. Pair index
    idx = p1 (p1 - 1) / 2 + p2, p1 > p2. Returns [(p1, p2)] in match order."""
    c = np.array(cost, np.float32)
    n = len(c)
    P = 1
    while P * (P - 1) // 2 < n:
        P += 1
    tau = np.float32(tau)
    out = []
    while True:
        vals = np.full(256, tau, np.float32)
        idxs = np.full(256, -1, np.int64)
        for tid in range(256):
            for i in range(tid, n, 256):
                if c[i] < vals[tid]:
                    vals[tid], idxs[tid] = c[i], i
        off = 128
        while off > 0:
            for tid in range(off):
                if vals[tid + off] < vals[tid]:
                    vals[tid], idxs[tid] = vals[tid + off], idxs[tid + off]
            off >>= 1
        bi = int(idxs[0])
        if bi < 0 or vals[0] >= tau:
            break
        p1 = 1
        while p1 * (p1 + 1) // 2 <= bi:
            p1 += 1
        p2 = bi - p1 * (p1 - 1) // 2
        out.append((p1, p2))
        c[bi] = MERGE_INVALID
        for k in range(P):
            for a, b in ((p1, k), (p2, k), (k, p1), (k, p2)):
                if a > b:
                    c[a * (a - 1) // 2 + b] = MERGE_INVALID
    return out


def merge_pass(parts, tau):
    """The merge-hulls post-pass, one round (csrc/lookahead.c:988-1067;
    cuda/postprocess_merge.cu): per pair (p1 > p2) of the same connected
    component (or unassigned), reject when the rv cost of the summed hull
    volumes is >= tau or the boxes are >= tau / 2 apart; else the exact hull
    of both hulls' vertices gives the merged rv cost, and below tau the
    Hausdorff distance between that hull and both meshes (triangles on a
    shared hull face dropped) raises it to max(rv, H). Matched pairs become
    one part in slot p1 (A's vertices and triangles, then B's), the other slot
    is removed, the order of the rest is kept."""
    f = np.float32
    tau = f(tau)
    P = len(parts)
    if P < 2:
        return parts
    cost = []
    hulls = {}
    for p1 in range(1, P):
        for p2 in range(p1):
            A, B = parts[p1], parts[p2]
            if len(A.v) == 0 or len(B.v) == 0 or A.hv is None or B.hv is None:
                cost.append(MERGE_INVALID)
                continue
            if A.cc != 0 and B.cc != 0 and A.cc != B.cc:
                cost.append(MERGE_INVALID)
                continue
            msum = A.mvol + B.mvol
            if cost_rv(A.hvol + B.hvol, msum) >= tau:
                cost.append(MERGE_RV_REJECT)
                continue
            g = [max(f(0), max(f(A.lo[j] - B.hi[j]), f(B.lo[j] - A.hi[j]))) for j in range(3)]
            g2 = f(f(f(g[0] * g[0]) + f(g[1] * g[1])) + f(g[2] * g[2]))
            half = f(tau * f(0.5))
            if g2 >= f(half * half):
                cost.append(MERGE_BBOX_REJECT)
                continue
            hv, ht, hvol = hull_mesh(np.concatenate([A.hv, B.hv]))
            if len(ht) == 0:
                cost.append(MERGE_INVALID)
                continue
            rv = cost_rv(hvol, msum)
            if rv >= tau:
                cost.append(rv)
                continue
            pl = shared_plane(A.hv, A.ht, B.hv)
            cv = np.concatenate([A.v, B.v])
            ct = np.concatenate([A.t, B.t + len(A.v)]).astype(np.int32)
            if pl is not None:
                nx, ny, nz, d = pl
                s = np.abs(((nx * cv[:, 0] + ny * cv[:, 1]) + nz * cv[:, 2]) - d) <= f(1e-3)
                ct = ct[~s[ct].all(axis=1)]
            h = hausdorff(hv, ht, cv, ct) if len(ct) else f(0)
            c = max(rv, f(h))
            cost.append(c)
            if c < tau:
                hulls[(p1, p2)] = (hv, ht, hvol)
    out = list(parts)
    dead = set()
    for p1, p2 in merge_match(cost, tau):
        A, B = parts[p1], parts[p2]
        hv, ht, hvol = hulls[(p1, p2)]
        m = Part(np.concatenate([A.v, B.v]), np.concatenate([A.t, B.t + len(A.v)]).astype(np.int32),
                 A.mvol + B.mvol, hvol, hv, ht, cc=A.cc if A.cc else B.cc)
        out[p1] = m
        dead.add(p2)
    return [q for i, q in enumerate(out) if i not in dead]



def lookahead(verts, tris, hull_v, hull_t, max_iters=100, width=60, width2=5, threshold=0.05, depth=2,
              max_n_cutting=16, decompose_components=False, dc_per_iter=True, n_concave_edges=32,
              concave_eps=0.005, concave_threshold=3.49, concave_iters=10, deadline=None, quick_depth=0,
              merge_hulls=False, trace=None):
    """Returns (parts, iters): parts = [(v, t, hv, ht, mesh_vol, hull_vol, hausdorff)].
    deadline (time.perf_counter() value): stop before an iteration that would
    start after it (bounded CPU-baseline samples). quick_depth: quick levels
    after the full ones (csrc/lookahead.c:656-727). trace: a list that gets
    one record per cutting part and iteration (its position after the sort,
    the best path cost, the chosen first cut's index and plane, and every
    first cut's best path cost before the greedy fallback)."""
    import time
    tau = np.float32(threshold)
    v0 = np.asarray(verts, np.float32)
    t0 = np.asarray(tris, np.int32)
    p0 = Part(v0, t0, mesh_volume_abs(v0, t0), mesh_volume_abs(hull_v, hull_t),
              np.asarray(hull_v, np.float32), np.asarray(hull_t, np.int32))
    parts = [p0]
    cc_counter = [0]
    med = np.float32(tau * np.float32(0.25))
    if med > np.float32(0.015):
        med = np.float32(0.015)
    total = depth + quick_depth  # csrc/lookahead.c:510

    def run_components():
        nonlocal parts
        new_tail = []
        for i in range(len(parts)):
            p = parts[i]
            if len(p.v) == 0 or len(p.t) == 0:
                continue
            comps = components(p.v, p.t)
            if comps is None:
                if p.cc == 0:
                    cc_counter[0] += 1
                    p.cc = cc_counter[0]
                continue
            fresh = p.cc == 0
            base = cc_counter[0]
            if fresh:
                cc_counter[0] += len(comps)
            for c, (cv, ct, sv) in enumerate(comps):
                q = Part(cv, ct, abs(sv), 0.0, cc=(base + c + 1) if fresh else p.cc)
                if c == 0:
                    parts[i] = q
                else:
                    new_tail.append(q)
        parts += new_tail

    def run_hulls():
        for p in parts:
            if p.hvol > 0.0 or len(p.v) == 0:
                continue
            p.hv, p.ht, p.hvol = hull_mesh(p.v)
            if p.mvol > p.hvol:
                p.mvol = p.hvol

    iters = 0
    for it in range(max_iters):
        if deadline is not None and time.perf_counter() > deadline:
            break
        if dc_per_iter:
            run_components()
        run_hulls()
        for p in parts:
            if p.hv is not None and len(p.ht) and p.haus <= 0 and cost_rv(p.hvol, p.mvol) < tau:
                p.haus = hausdorff(p.hv, p.ht, p.v, p.t)
        order, cutting = select_cutting([q.cost() for q in parts], tau, max_n_cutting)
        parts = [parts[i] for i in order]
        if not cutting:
            break
        iters = it + 1
        eplanes = [[] for _ in cutting]
        if n_concave_edges > 0 and it < concave_iters:
            thr = float(np.float32(concave_threshold))
            eplanes = [concave_planes(parts[i], n_concave_edges, concave_eps, thr) for i in cutting]
        n_edge = max(len(e) for e in eplanes) if eplanes else 0
        extra = []
        l0 = [dict() for _ in cutting]
        l0_planes = [dict() for _ in cutting]
        # items: parts as (mvol, hvol, v, t, lo, hi) tuples, worst last
        cur = []
        for c, i in enumerate(cutting):
            p = parts[i]
            cur.append(Item(c, 0, [(p.mvol, p.hvol, p.v, p.t, p.lo, p.hi)], []))
        leaves = []
        leaf_recs = []
        for d in range(depth):
            last = d == depth - 1
            # the last full level is only scored (branch-and-bound, no caps)
            # unless quick levels follow (csrc/lookahead.c:589)
            full = (not last) or depth == 1 or quick_depth > 0
            w = width if d == 0 else width2
            nxt = []
            bub = np.float32(1e30) if full else best_ub([itm.lc for itm in cur], total)
            any_child = False
            for itm in cur:
                tw = w + (n_edge if d == 0 else 0)
                wm, wh, wv, wt, wlo, whi = itm.parts[-1]
                assert len(itm.parts) + 1 <= LA_MAX_PARTS
                for k in range(tw):
                    if k >= w:
                        ep = eplanes[itm.src]
                        e = k - w
                        pl = ep[e] if e < len(ep) else np.zeros(4, np.float32)
                        if pl[0] == 0 and pl[1] == 0 and pl[2] == 0:
                            continue
                    ext = np.float32(np.float32(2.0) * med + np.float32(1e-6))
                    small = all(np.float32(whi[j] - wlo[j]) <= ext for j in range(3))
                    if small:
                        if k == 0:
                            lc = list(itm.lc) + [np.float32(0)] * (min(total, LA_MAX_LEVELS) - len(itm.lc))
                            extra.append(Item(itm.src, itm.cut0, itm.parts, lc))
                        continue
                    if k < w:
                        ax, sl, cpa = axis_cut(w, k)
                        lov = np.float32(wlo[ax] + med)
                        hiv = np.float32(whi[ax] - med)
                        if lov >= hiv:
                            continue
                        frac = np.float32(np.float32(sl + 1.0) / np.float32(cpa + 1.0))
                        pos = np.float32(lov + np.float32(frac * np.float32(hiv - lov)))
                        pl = np.zeros(4, np.float32)
                        pl[ax] = 1.0
                        pl[3] = -pos
                    cut0 = k if len(itm.lc) == 0 else itm.cut0
                    if full:
                        st, pv, pt, nv, nt = O.plane_cut(wv, wt, pl)
                        if st != 0 or len(pv) == 0 or len(nv) == 0:
                            continue
                        pm, nmv = mesh_volume_abs(pv, pt), mesh_volume_abs(nv, nt)
                        ph, _ = hull_volume(pv)
                        nh, _ = hull_volume(nv)
                        ch = [x for x in itm.parts[:-1]] + [(pm, ph, pv, pt, pv.min(0), pv.max(0)),
                                                            (nmv, nh, nv, nt, nv.min(0), nv.max(0))]
                        ch = _sort_parts(ch)
                        lc = list(itm.lc)
                        if len(lc) < LA_MAX_LEVELS:
                            lc.append(cost_rv(ch[-1][1], ch[-1][0]))
                        c_item = Item(itm.src, cut0, ch, lc,
                                      store=(Part(pv, pt, pm), Part(nv, nt, nmv)) if d == 0 else None)
                        nxt.append(c_item)
                        any_child = True
                        if d == 0:
                            l0[itm.src][k] = c_item
                            l0_planes[itm.src][k] = np.array(pl, np.float32)
                    else:
                        pvol, nvol, npc, nnc, ppts, npts = score_cut(wv, wt, pl)
                        if npc == 0 or nnc == 0:
                            continue
                        gap = prune_gap(itm.lc, bub, total)
                        ph, prp = hull_volume(ppts, gap, pvol)
                        nh, prn = hull_volume(npts, gap, nvol)
                        worst = max(cost_rv(ph, pvol), cost_rv(nh, nvol))
                        for x in itm.parts[:-1]:
                            worst = max(worst, cost_rv(x[1], x[0]))
                        leaf_recs.append((itm.src, cut0, path_cost(list(itm.lc) + [worst]), prp or prn))
                        any_child = True
            if not full:
                if not any_child:
                    leaves.extend(cur)
                cur = []
                break
            if not any_child:
                break  # the items stay (csrc/lookahead.c:579-581)
            cur = nxt
        cur = quick_levels(cur, extra, quick_depth, total, med)
        leaves.extend(cur)
        # evaluate + apply
        tw0 = width + n_edge
        paths = [[] for _ in cutting]
        for itm in leaves:
            if not itm.pruned and len(itm.lc) > 0:
                paths[itm.src].append((itm.cut0, path_cost(itm.lc)))
        for src, cut0, path, pruned in leaf_recs:
            if not pruned:
                paths[src].append((cut0, path))
        for itm in extra:
            paths[itm.src].append((itm.cut0, path_cost(itm.lc)))
        new_parts = []
        for c, i in enumerate(cutting):
            l0c = {k: itm.lc[0] for k, itm in l0[c].items() if len(itm.lc) >= 1}
            bc, bi = evaluate(tw0, paths[c], l0c)
            if trace is not None:
                trace.append({"iter": it, "part": i, "best_cost": bc, "best_cut": bi,
                              "plane": l0_planes[c].get(bi), "per_cut": per_cut_best(tw0, paths[c])})
            if bi is None or bi not in l0[c]:
                continue
            pos, neg = l0[c][bi].store
            cc = parts[i].cc
            pos.cc = neg.cc = cc
            parts[i] = pos
            new_parts.append(neg)
        parts += new_parts
    else:
        pass
    if iters == max_iters:
        if dc_per_iter:
            run_components()
        run_hulls()
    if decompose_components:
        run_components()
        run_hulls()
    if merge_hulls:
        parts = merge_pass(parts, tau)
    return [(p.v, p.t, p.hv, p.ht, p.mvol, p.hvol, p.haus) for p in parts], iters
