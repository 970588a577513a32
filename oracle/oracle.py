"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product path (cuacd_amd).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run python cuacd_amd/build.py oracle")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_quantize.argtypes = [P, C.c_int, P]
        L.orc_hull.argtypes = [P, C.c_int, P, P, C.c_int]
        L.orc_hull.restype = C.c_int
        L.orc_hull_canon.argtypes = [P, C.c_int, P, P, C.c_int]
        L.orc_hull_canon.restype = C.c_int
        L.orc_general_position.argtypes = [P, C.c_int, P, C.c_int]
        L.orc_general_position.restype = C.c_int
        L.orc_mesh_volume.argtypes = [P, P, C.c_int]
        L.orc_mesh_volume.restype = C.c_double
        L.orc_hull_volume.argtypes = [P, P, C.c_int]
        L.orc_hull_volume.restype = C.c_double
        L.orc_hull_bruteforce.argtypes = [P, C.c_int, P]
        L.orc_hull_bruteforce.restype = C.c_int
        L.orc_hull_batch.argtypes = [P, P, C.c_int, P, P, P, P, C.c_int]
        L.orc_hull_batch.restype = C.c_int
        L.orc_hull_batch_serial.argtypes = [P, P, C.c_int, P, P, P]
        L.orc_hull_batch_serial.restype = C.c_int
        L.orc_pt_tri_d2.argtypes = [P, P, P, P]
        L.orc_pt_tri_d2.restype = C.c_double
        L.orc_points_mesh_d2.argtypes = [P, C.c_int, P, P, C.c_int, P]
        L.orc_directed_hausdorff.argtypes = [P, C.c_int, P, P, C.c_int]
        L.orc_directed_hausdorff.restype = C.c_double
        L.orc_directed_batch.argtypes = [P, C.c_int, P, P, P, P, C.c_int, P]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _LIB = L
    return _LIB


def set_threads(n):
    """OpenMP threads of the batch functions (hull_batch, directed_batch, ...)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def quantize(pts):
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    q = np.empty((len(pts), 3), np.int32)
    lib().orc_quantize(_p(pts), len(pts), _p(q))
    return q


def hull(q, canonical=True):
    """Exact hull of integer points: (vertex flags bool[n], faces int[nf,3]).
    canonical: corners only, facets as fans from their lowest-index corner,
    ordered by (c0, c1) (the form the GPU library emits); otherwise the
    incremental construction's own faces and vertices."""
    q = np.ascontiguousarray(q, np.int32).reshape(-1, 3)
    n = len(q)
    flags = np.zeros(max(n, 1), np.uint8)
    cap = 2 * n + 8
    faces = np.zeros((cap, 3), np.int32)
    fn = lib().orc_hull_canon if canonical else lib().orc_hull
    nf = fn(_p(q), n, _p(flags), _p(faces), cap)
    if nf < 0:
        raise RuntimeError("oracle hull failed")
    return flags[:n].astype(bool), faces[:nf].copy()


def general_position(q, faces):
    q = np.ascontiguousarray(q, np.int32)
    faces = np.ascontiguousarray(faces, np.int32)
    return bool(lib().orc_general_position(_p(q), len(q), _p(faces), len(faces)))


def mesh_volume(pts, tris):
    pts = np.ascontiguousarray(pts, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    return lib().orc_mesh_volume(_p(pts), _p(tris), len(tris))


def hull_volume(pts, tris):
    """FP64 volume of a canonical hull mesh in the GPU library's fixed
    summation order (64 contiguous chunks, then their partial sums)."""
    pts = np.ascontiguousarray(pts, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    return lib().orc_hull_volume(_p(pts), _p(tris), len(tris))


def hull_bruteforce(q):
    q = np.ascontiguousarray(q, np.int32).reshape(-1, 3)
    flags = np.zeros(max(len(q), 1), np.uint8)
    nf = lib().orc_hull_bruteforce(_p(q), len(q), _p(flags))
    return flags[:len(q)].astype(bool), nf


def hull_batch(pts, offsets, check_general=True, serial=False):
    """Per cloud: vertex flags (concatenated), face counts, FP64 volumes, general flags."""
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    offsets = np.ascontiguousarray(offsets, np.int32)
    n = len(offsets) - 1
    flags = np.zeros(max(len(pts), 1), np.uint8)
    nf = np.zeros(n, np.int32)
    vol = np.zeros(n, np.float64)
    gen = np.zeros(n, np.int32)
    if serial:
        bad = lib().orc_hull_batch_serial(_p(pts), _p(offsets), n, _p(flags), _p(nf), _p(vol))
    else:
        bad = lib().orc_hull_batch(_p(pts), _p(offsets), n, _p(flags), _p(nf), _p(vol), _p(gen),
                                   1 if check_general else 0)
    if bad:
        raise RuntimeError(f"oracle hull failed on {bad} clouds")
    return flags[:len(pts)].astype(bool), nf, vol, gen


def pt_tri_d2(p, a, b, c):
    arr = [np.ascontiguousarray(x, np.float64) for x in (p, a, b, c)]
    return lib().orc_pt_tri_d2(*[_p(x) for x in arr])


def points_mesh_d2(pts, verts, tris):
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    out = np.empty(len(pts), np.float64)
    lib().orc_points_mesh_d2(_p(pts), len(pts), _p(verts), _p(tris), len(tris), _p(out))
    return out


def directed_hausdorff(pts, verts, tris):
    """max over pts of the exact distance to the mesh surface (FP64 brute force)."""
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    return lib().orc_directed_hausdorff(_p(pts), len(pts), _p(verts), _p(tris), len(tris))


def directed_batch(samples, meshes):
    """samples: (n_pairs, n_samples, 3) float32; meshes: list of (verts, tris)."""
    samples = np.ascontiguousarray(samples, np.float32)
    n = len(meshes)
    vs = [np.ascontiguousarray(v, np.float32).reshape(-1, 3) for v, _ in meshes]
    ts = [np.ascontiguousarray(t, np.int32).reshape(-1, 3) for _, t in meshes]
    voff = np.zeros(n + 1, np.int32)
    toff = np.zeros(n + 1, np.int32)
    voff[1:] = np.cumsum([len(v) for v in vs])
    toff[1:] = np.cumsum([len(t) for t in ts])
    av = np.ascontiguousarray(np.concatenate(vs))
    at = np.ascontiguousarray(np.concatenate(ts))
    out = np.empty(n, np.float64)
    lib().orc_directed_batch(_p(samples), samples.shape[1], _p(av), _p(voff), _p(at), _p(toff), n, _p(out))
    return out


def reference_sampler_counts(verts, tris):
    """Per-triangle sample counts of the reference Hausdorff sampler
    (cuda/hausdorff.cuh:380-382, 432-466), evaluated in float32 except for the
    total area, an FP64 sum of the float32 triangle areas rounded once (the
    reference sums in float32 in an order of its own; a rounded FP64 sum is
    what a GPU reduction and a serial loop can both reproduce)."""
    v = np.asarray(verts, np.float32)
    t = np.asarray(tris, np.int64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    n = np.cross(b - a, c - a).astype(np.float32)
    area_t = (np.float32(0.5) * np.sqrt((n * n).sum(1, dtype=np.float32))).astype(np.float32)
    area = np.float32(area_t.astype(np.float64).sum())  # FP64 sum, rounded once
    res = max(int(np.float32(2000.0) * area), 1000)
    nt = len(t)
    step = nt // res if nt > res else 2
    dens = np.float32(res) / area if area > 1e-20 else np.float32(0)
    ac = (dens * area_t).astype(np.float32).astype(np.int64)
    idx = np.arange(nt)
    return np.where(ac > 0, ac, (idx % step == 0).astype(np.int64)), res


def _wang(s):
    s = np.asarray(s, np.uint64) & np.uint64(0xFFFFFFFF)
    s = (s ^ np.uint64(61)) ^ (s >> np.uint64(16))
    s = (s * np.uint64(9)) & np.uint64(0xFFFFFFFF)
    s = s ^ (s >> np.uint64(4))
    s = (s * np.uint64(0x27D4EB2D)) & np.uint64(0xFFFFFFFF)
    s = s ^ (s >> np.uint64(15))
    return s


def reference_samples(verts, tris):
    """Sample points of the reference Hausdorff sampler (cuda/hausdorff.cuh:516-558):
    triangle t gets counts[t] points, point k from Wang hashes of t*65537+k,
    barycentric (1-sqrt(ra), sqrt(ra)(1-rb), sqrt(ra) rb)."""
    v = np.asarray(verts, np.float32)
    t = np.asarray(tris, np.int64)
    counts, _ = reference_sampler_counts(v, t)
    tri = np.repeat(np.arange(len(t)), counts)
    k = np.concatenate([np.arange(c) for c in counts]) if len(counts) else np.zeros(0, np.int64)
    ha = _wang(tri.astype(np.uint64) * np.uint64(65537) + k.astype(np.uint64))
    hb = _wang(ha + np.uint64(0x9E3779B9))
    ra = (ha & np.uint64(0xFFFFFF)).astype(np.float32) / np.float32(0xFFFFFF)
    rb = (hb & np.uint64(0xFFFFFF)).astype(np.float32) / np.float32(0xFFFFFF)
    sqa = np.sqrt(ra)
    a, b, c = v[t[tri, 0]], v[t[tri, 1]], v[t[tri, 2]]
    w = np.stack([1 - sqa, sqa * (1 - rb), sqa * rb], 1).astype(np.float32)
    return (w[:, :1] * a + w[:, 1:2] * b + w[:, 2:3] * c).astype(np.float32)


def bidirectional_hausdorff_reference(va, ta, vb, tb):
    """sqrt(max(dA->B, dB->A)) with reference-sampler points and exact distances."""
    sa = reference_samples(va, ta)
    sb = reference_samples(vb, tb)
    return max(directed_hausdorff(sa, vb, tb), directed_hausdorff(sb, va, ta))


def halfspace_volumes(verts, tris, plane):
    """FP64 volumes of the mesh on each side of plane (a, b, c, d): (pos, neg)."""
    L = lib()
    if not getattr(L, "_hs", False):
        L.orc_halfspace_volumes.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                            C.c_double, C.c_double, C.c_void_p]
        L._hs = True
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.int32)
    out = np.zeros(2, np.float64)
    L.orc_halfspace_volumes(_p(v), _p(t), len(t), *[float(x) for x in plane], _p(out))
    return float(out[0]), float(out[1])


def edge_manifold_report(tris):
    """(open directed edges, duplicated directed edges) - 0, 0 for a closed,
    consistently oriented 2-manifold (audit semantics of cuda/mesh_audit.cuh)."""
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    key = e[:, 0] * (1 << 32) + e[:, 1]
    twin = e[:, 1] * (1 << 32) + e[:, 0]
    uk, cnt = np.unique(key, return_counts=True)
    dup = int((cnt > 1).sum())
    open_ = int((~np.isin(twin, uk)).sum())
    return open_, dup


def plane_cut(verts, tris, plane):
    """Serial plane cut with caps: (status, pos_v, pos_t, neg_v, neg_t)."""
    L = lib()
    if not getattr(L, "_pc", False):
        L.orc_plane_cut.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int] + [C.c_float] * 4 + \
            [C.c_void_p] * 4 + [C.c_int, C.c_int, C.c_void_p]
        L.orc_plane_cut.restype = C.c_int
        L._pc = True
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    cap_v = len(v) + 3 * len(t) + 16
    cap_t = 3 * len(t) + 2 * cap_v + 16
    pv = np.zeros((cap_v, 3), np.float32)
    pt = np.zeros((cap_t, 3), np.int32)
    nv_ = np.zeros((cap_v, 3), np.float32)
    nt_ = np.zeros((cap_t, 3), np.int32)
    sz = np.zeros(4, np.int32)
    a, b, c, d = (float(np.float32(x)) for x in plane)
    st = L.orc_plane_cut(_p(v), len(v), _p(t), len(t), a, b, c, d, _p(pv), _p(pt), _p(nv_), _p(nt_),
                         cap_v, cap_t, _p(sz))
    if st < 0:
        raise RuntimeError("orc_plane_cut: capacity")
    return st, pv[:sz[0]].copy(), pt[:sz[1]].copy(), nv_[:sz[2]].copy(), nt_[:sz[3]].copy()
