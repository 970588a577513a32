"""Seeded SDF solids and a watertight iso-surface mesher (no mesh libraries).

The benchmark and test meshes are iso-surfaces of random signed distance
fields: a smooth union of 3-8 random capsules and boxes minus 0-3 spheres,
sampled on an R^3 grid over [-1.2, 1.2]^3 (R = 48 by default: marching
tetrahedra emit about twice the triangles of marching cubes, so 48^3 gives the
triangle counts of marching cubes at 64^3) and polygonised by marching
tetrahedra (each grid cube split into the 6 Kuhn tetrahedra that share its
main diagonal, so neighbouring cubes agree on every face diagonal). Every
crossing edge gets one shared vertex, so the surface is closed and 2-manifold
whenever no grid value is exactly zero (zeros are nudged). Triangles are
oriented so that their normals point towards increasing SDF (outwards). The
mesh is then centred and scaled to extent 2 (the largest bbox side).
"""
import numpy as np

# Kuhn tetrahedra: paths 0 -> e_a -> e_a + e_b -> (1,1,1) over axis permutations.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _tets():
    out = []
    for p in _PERMS:
        c = [np.zeros(3, int)]
        for ax in p:
            n = c[-1].copy()
            n[ax] = 1
            c.append(n)
        out.append(np.array(c))
    return out  # 6 x [4, 3] corner offsets


def random_sdf_params(rng):
    """Primitive list for one solid: ('capsule', a, b, r) / ('box', c, half, R) /
    ('sphere', c, r) subtractions, plus the smoothing radius."""
    n_add = int(rng.integers(3, 9))
    n_sub = int(rng.integers(0, 4))
    prims = []
    for _ in range(n_add):
        if rng.random() < 0.5:
            a = rng.uniform(-0.7, 0.7, 3)
            b = a + rng.normal(0, 0.5, 3)
            prims.append(("capsule", a, np.clip(b, -0.9, 0.9), float(rng.uniform(0.08, 0.3))))
        else:
            c = rng.uniform(-0.6, 0.6, 3)
            half = rng.uniform(0.1, 0.45, 3)
            q, r = np.linalg.qr(rng.standard_normal((3, 3)))
            prims.append(("box", c, half, q * np.sign(np.diag(r))))
    subs = [("sphere", rng.uniform(-0.6, 0.6, 3), float(rng.uniform(0.1, 0.35))) for _ in range(n_sub)]
    return prims, subs, float(rng.uniform(0.03, 0.12))


def _smin(a, b, k):
    h = np.clip(0.5 + 0.5 * (b - a) / k, 0.0, 1.0)
    return b + (a - b) * h - k * h * (1.0 - h)


def eval_sdf(params, p):
    """SDF at points p [..., 3] (float64)."""
    prims, subs, k = params
    d = None
    for pr in prims:
        if pr[0] == "capsule":
            _, a, b, r = pr
            ab = b - a
            t = np.clip(((p - a) @ ab) / max(ab @ ab, 1e-12), 0.0, 1.0)
            di = np.linalg.norm(p - (a + t[..., None] * ab), axis=-1) - r
        else:
            _, c, half, R = pr
            q = np.abs((p - c) @ R) - half
            di = np.linalg.norm(np.maximum(q, 0.0), axis=-1) + np.minimum(q.max(-1), 0.0)
        d = di if d is None else _smin(d, di, k)
    for _, c, r in subs:
        ds = np.linalg.norm(p - c, axis=-1) - r
        d = np.maximum(d, -ds)
    return d


def marching_tets(f, lo, h):
    """Iso-surface f = 0 of grid values f [R, R, R] at points lo + h * (i, j, k).
    Returns float64 verts [V, 3], int32 tris [T, 3]."""
    f = np.array(f, np.float64)
    f[f == 0.0] = 1e-9
    R = np.array(f.shape)
    ii = np.stack(np.meshgrid(*[np.arange(n - 1) for n in R], indexing="ij"), -1).reshape(-1, 3)
    lin = lambda c: (c[:, 0] * R[1] + c[:, 1]) * R[2] + c[:, 2]
    fl = f.reshape(-1)
    edge_key, edge_t, tris_e = [], [], []
    for tet in _tets():
        corners = [lin(ii + o) for o in tet]          # 4 x [n_cubes] global corner ids
        vals = np.stack([fl[c] for c in corners], 1)  # [n, 4]
        inside = vals < 0
        cnt = inside.sum(1)
        for want in (1, 2, 3):
            sel = np.nonzero(cnt == want)[0]
            if len(sel) == 0:
                continue
            ins = inside[sel]
            order = np.argsort(~ins, axis=1, kind="stable")   # inside corners first
            cid = np.stack(corners, 1)[sel]
            oc = np.take_along_axis(cid, order, 1)
            ov = np.take_along_axis(vals[sel], order, 1)
            if want == 1:
                pairs = [[(0, 1), (0, 2), (0, 3)]]
            elif want == 3:
                pairs = [[(0, 3), (1, 3), (2, 3)]]
            else:
                pairs = [[(0, 2), (0, 3), (1, 3)], [(0, 2), (1, 3), (1, 2)]]
            for tri in pairs:
                ks = []
                for a, b in tri:
                    ga, gb = oc[:, a], oc[:, b]
                    va, vb = ov[:, a], ov[:, b]
                    sw = ga > gb
                    k0 = np.where(sw, gb, ga)
                    k1 = np.where(sw, ga, gb)
                    f0 = np.where(sw, vb, va)
                    f1 = np.where(sw, va, vb)
                    edge_key.append(k0.astype(np.int64) * fl.size + k1)
                    edge_t.append(f0 / (f0 - f1))
                    ks.append(len(edge_key) - 1)
                # inside/outside centroids fix the orientation later
                cin = oc[:, :want]
                cout = oc[:, want:]
                tris_e.append((ks, cin, cout))
    keys = np.concatenate(edge_key)
    ts = np.concatenate(edge_t)
    uk, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    k0, k1 = uk // fl.size, uk % fl.size
    t = ts[first]
    coord = lambda g: np.stack([g // (R[1] * R[2]), (g // R[2]) % R[1], g % R[2]], -1).astype(np.float64)
    verts = lo + h * (coord(k0) + t[:, None] * (coord(k1) - coord(k0)))
    offs = np.cumsum([0] + [len(k) for k in edge_key])
    all_tris = []
    for ks, cin, cout in tris_e:
        idx = np.stack([inv[offs[k]:offs[k + 1]] for k in ks], 1)
        p = verts[idx]
        nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        gi = coord(cin.reshape(-1)).reshape(cin.shape + (3,)).mean(1)
        go = coord(cout.reshape(-1)).reshape(cout.shape + (3,)).mean(1)
        flip = (nrm * (go - gi)).sum(1) < 0
        idx[flip] = idx[flip][:, [0, 2, 1]]
        all_tris.append(idx)
    tris = np.concatenate(all_tris).astype(np.int32)
    # drop triangles made degenerate by coincident vertices (t exactly 0 or 1)
    good = (tris[:, 0] != tris[:, 1]) & (tris[:, 1] != tris[:, 2]) & (tris[:, 0] != tris[:, 2])
    return verts, tris[good]


def normalize_extent2(v):
    lo, hi = v.min(0), v.max(0)
    c = (lo + hi) / 2
    ext = (hi - lo).max()
    return (v - c) * (2.0 / ext if ext > 0 else 1.0)


def sdf_mesh(params, res=48):
    g = np.linspace(-1.2, 1.2, res)
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    f = eval_sdf(params, P)
    # a positive shell keeps solids that reach the grid border closed
    f[0, :, :] = f[-1, :, :] = f[:, 0, :] = f[:, -1, :] = f[:, :, 0] = f[:, :, -1] = 1.0
    v, t = marching_tets(f, np.array([-1.2, -1.2, -1.2]), g[1] - g[0])
    return normalize_extent2(v).astype(np.float32), t


def sdf_meshes(n, seed=0, res=48, min_tris=0, max_tris=None):
    """n seeded SDF meshes; solids outside [min_tris, max_tris] are redrawn
    from the same stream."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        v, t = sdf_mesh(random_sdf_params(rng), res)
        if len(t) < max(min_tris, 4) or (max_tris is not None and len(t) > max_tris):
            continue
        out.append((v, t))
    return out
