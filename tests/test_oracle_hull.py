"""CPU tests that pin the hull oracle (oracle/hull_oracle.c) before it judges the GPU.

Pins: the reference's known answers (cube 8 verts/12 tris, tetrahedron 4/4,
unit-cube volume 1: tests/test_hull_mesh.py:90-120), brute-force facet
enumeration on small clouds, scipy Qhull vertex sets on the same integer points,
hand-computed quantization, and closed-form volume properties (scaling, translation).
"""
import numpy as np
import pytest
import scipy.spatial

import oracle
from cuacd_amd import datasets


def _cube():
    return np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)


def test_quantize_closed_form():
    # AABB (0..2, 0..1, 0..0.5): max axis x, min axis z, median y.
    # (med+1)%3 = 2 != max=0 -> scaling negated. q = (y-0.5)*-10216, (x-1)*-5108, (z-.25)*-20432
    pts = np.array([[0, 0, 0], [2, 1, 0.5], [1.5, 0.25, 0.125]], np.float32)
    q = oracle.quantize(pts)
    assert q.tolist() == [[5108, 5108, 5108], [-5108, -5108, -5108], [2554, -2554, 2554]]


def test_quantize_right_handed_axes():
    # max axis y, min axis z, median x: (0+1)%3 == 1 == max -> no sign flip
    pts = np.array([[0, 0, 0], [1, 2, 0.5]], np.float32)
    q = oracle.quantize(pts)
    assert q.tolist() == [[-5108, -5108, -5108], [5108, 5108, 5108]]


def test_cube_and_tetrahedron_known_answers():
    q = oracle.quantize(_cube())
    flags, faces = oracle.hull(q)
    assert flags.sum() == 8 and len(faces) == 12
    assert abs(oracle.mesh_volume(_cube(), faces) - 1.0) < 1e-12
    tet = np.array([[1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]], np.float32)
    flags, faces = oracle.hull(oracle.quantize(tet))
    assert flags.sum() == 4 and len(faces) == 4
    assert abs(oracle.mesh_volume(tet, faces) - 8.0 / 3.0) < 1e-12


def test_faces_outward_and_closed():
    pts = datasets.ball_clouds(4, 200, seed=7)
    for p in pts:
        q = oracle.quantize(p)
        flags, faces = oracle.hull(q)
        # every directed edge appears exactly once and its twin exactly once
        edges = set()
        for a, b, c in faces:
            for e in ((a, b), (b, c), (c, a)):
                assert e not in edges
                edges.add(e)
        for a, b in edges:
            assert (b, a) in edges
        # positive volume = outward winding
        assert oracle.mesh_volume(p, faces) > 0
        # Euler for a simplicial sphere
        assert len(faces) == 2 * flags.sum() - 4


@pytest.mark.parametrize("n", [5, 8, 13, 21, 34])
def test_matches_bruteforce(n):
    rng = np.random.default_rng(n)
    done = 0
    for _ in range(40):
        p = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        q = oracle.quantize(p)
        flags, faces = oracle.hull(q)
        if not oracle.general_position(q, faces):
            continue
        bf, nfac = oracle.hull_bruteforce(q)
        assert np.array_equal(flags, bf)
        assert nfac == len(faces)
        done += 1
    assert done >= 30


def test_matches_qhull_on_integer_grid():
    clouds = datasets.ball_clouds(64, 256, seed=0)
    for p in clouds:
        q = oracle.quantize(p)
        flags, faces = oracle.hull(q)
        if not oracle.general_position(q, faces):
            continue
        qh = scipy.spatial.ConvexHull(q.astype(np.float64))
        ref = np.zeros(len(q), bool)
        ref[qh.vertices] = True
        assert np.array_equal(flags, ref)


def test_duplicates_resolve_to_lowest_index():
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, (60, 3)).astype(np.float32)
    # append exact copies of the first 30 points (higher indices)
    p2 = np.concatenate([p, p[:30]], 0)
    flags1, _ = oracle.hull(oracle.quantize(p))
    flags2, _ = oracle.hull(oracle.quantize(p2))
    assert np.array_equal(flags2[:60], flags1)
    assert not flags2[60:].any()
    # copies placed BEFORE the originals win instead
    p3 = np.concatenate([p[:30], p], 0)
    flags3, _ = oracle.hull(oracle.quantize(p3))
    assert np.array_equal(flags3[:30], flags1[:30])
    assert not flags3[30:60].any()


def test_volume_properties():
    pts = datasets.ball_clouds(8, 256, seed=11)
    for p in pts:
        _, f = oracle.hull(oracle.quantize(p))
        v = oracle.mesh_volume(p, f)
        # sphere-ish: close to 4/3 pi but below it
        assert 2.7 < v < 4.0 * np.pi / 3.0
        # scaling by 2 (exact in fp32) multiplies the volume by 8; the grid
        # scales with the cloud, so the hull is the same
        p2 = p * np.float32(2.0)
        _, f2 = oracle.hull(oracle.quantize(p2))
        assert abs(oracle.mesh_volume(p2, f2) - 8 * v) < 1e-9 * v
        # volume agrees with Qhull's on the float points to quantization accuracy
        assert abs(scipy.spatial.ConvexHull(p.astype(np.float64)).volume - v) < 2e-3 * v


def test_batch_equals_single_and_reports_general_position():
    pts = datasets.ball_clouds(32, 256, seed=0)
    packed = np.concatenate(pts, 0)
    off = np.arange(33, dtype=np.int32) * 256
    flags, nf, vol, gen = oracle.hull_batch(packed, off)
    for i, p in enumerate(pts):
        q = oracle.quantize(p)
        f1, faces = oracle.hull(q)
        assert np.array_equal(flags[off[i]:off[i + 1]], f1)
        assert nf[i] == len(faces)
        assert vol[i] == oracle.hull_volume(p, faces)
        assert gen[i] == int(oracle.general_position(q, faces))
    assert gen.sum() >= 24  # the uniform-ball protocol is overwhelmingly in general position


def test_planted_mistake_is_caught():
    """A wrong sign in the oracle's orientation would flip every face; the
    closed-form checks above must catch that: emulate by reversing faces."""
    p = _cube()
    _, f = oracle.hull(oracle.quantize(p))
    assert oracle.mesh_volume(p, f) > 0
    assert oracle.mesh_volume(p, f[:, ::-1].copy()) < 0


# ---- canonical form (what the GPU library emits) -----------------------------

def _facets(q, tris):
    """Group triangles by exact supporting plane: {plane key: [triangles]}."""
    out = {}
    for t in tris:
        a, b, c = (q[i].astype(np.int64) for i in t)
        n = np.cross(b - a, c - a)
        g = np.gcd.reduce(np.abs(n))
        n = n // g
        out.setdefault((*n.tolist(), int(n @ a)), []).append(tuple(int(i) for i in t))
    return out


def _degenerate_cloud(rng, n=120, span=3):
    """Points on a tiny integer lattice: many coplanar and collinear points,
    duplicates included."""
    return rng.integers(0, span + 1, size=(n, 3)).astype(np.int32)


@pytest.mark.parametrize("seed", range(6))
def test_canonical_hull_on_degenerate_lattices(seed):
    """Corners = Qhull's extreme points; every facet a fan from its lowest-index
    corner with no degenerate triangle; all points on or below every facet;
    closed and outward; volume = Qhull's; facets ordered by (c0, c1)."""
    rng = np.random.default_rng(seed)
    q = _degenerate_cloud(rng, span=2 + seed % 3)
    flags, tris = oracle.hull(q)
    if len(tris) == 0:
        pytest.skip("flat sample")
    # duplicates resolve to the lowest index: compare corner coordinates
    qh = scipy.spatial.ConvexHull(q.astype(np.float64))
    want = sorted(set(map(tuple, q[qh.vertices].tolist())))
    got = sorted(map(tuple, q[flags].tolist()))
    assert got == want
    for i in np.nonzero(flags)[0]:
        assert not any((q[j] == q[i]).all() for j in range(i)), "a corner must be the lowest-index duplicate"
    assert oracle.edge_manifold_report(tris) == (0, 0)
    assert abs(oracle.mesh_volume(q.astype(np.float32), tris) - qh.volume) < 1e-9
    fac = _facets(q, tris)
    for (nx, ny, nz, d), ts in fac.items():
        nrm = np.array([nx, ny, nz], np.int64)
        assert (q.astype(np.int64) @ nrm <= d).all()
        c0 = ts[0][0]
        assert all(t[0] == c0 for t in ts)
        corners = {i for t in ts for i in t}
        assert c0 == min(corners)
        assert len(ts) == len(corners) - 2
        for t in ts:
            a, b, c = (q[i].astype(np.int64) for i in t)
            assert np.cross(b - a, c - a).any()
    # facets ordered by (c0, c1): triangles sorted by their first two vertices
    # at facet starts
    starts, seen = [], set()
    for t in tris:
        a, b, c = (q[i].astype(np.int64) for i in t)
        key = None
        for k, ts in fac.items():
            if tuple(int(i) for i in t) in ts:
                key = k
        if key not in seen:
            seen.add(key)
            starts.append((t[0], t[1]))
    assert starts == sorted(starts)


def test_canonical_hull_cube_with_face_and_edge_points():
    """A cube with extra points in face interiors and on edges: 8 corners,
    12 triangles, each square a 2-triangle fan from its lowest corner."""
    cube = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0],
                     [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2]], np.int32)
    extra = np.array([[1, 1, 0], [1, 0, 0], [2, 1, 1], [1, 2, 2], [0, 1, 1], [1, 1, 2]], np.int32)
    for order in (np.arange(14), np.random.default_rng(5).permutation(14)):
        q = np.concatenate([extra, cube])[order]
        flags, tris = oracle.hull(q)
        assert sorted(map(tuple, q[flags].tolist())) == sorted(map(tuple, cube.tolist()))
        assert len(tris) == 12
        assert abs(oracle.mesh_volume(q.astype(np.float32), tris) - 8.0) < 1e-12
        for ts in _facets(q, tris).values():
            corners = {i for t in ts for i in t}
            assert len(ts) == 2 and ts[0][0] == ts[1][0] == min(corners)


def test_canonical_hull_equals_raw_in_general_position():
    rng = np.random.default_rng(3)
    p = rng.standard_normal((300, 3)).astype(np.float32)
    q = oracle.quantize(p)
    f_raw, t_raw = oracle.hull(q, canonical=False)
    f_can, t_can = oracle.hull(q)
    assert oracle.general_position(q, t_raw)
    assert (f_raw == f_can).all()
    assert sorted(map(lambda t: tuple(np.roll(t, -int(np.argmin(t)))), t_raw.tolist())) == \
        sorted(map(tuple, t_can.tolist()))
    assert abs(oracle.mesh_volume(p, t_raw) - oracle.mesh_volume(p, t_can)) < 1e-12


def test_chunked_hull_union_keeps_the_canonical_hull():
    """The property the GPU's chunked path for oversized sets rests on: with
    one quantization frame, the canonical hull of the union of the chunks'
    canonical corners (chunks contiguous, corners in point order, a flat
    chunk passed through whole) is the canonical hull of the whole set."""
    rng = np.random.default_rng(8)
    p = rng.normal(size=(3000, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p = p[p[:, 0] >= 0.1] * (1 + 0.002 * rng.normal(size=(1, 1)))
    # a planar cap on x = 0.1 appended last, as a cut emits it
    cap = np.c_[np.full(400, 0.1), rng.uniform(-0.9, 0.9, (400, 2))]
    cap = cap[np.linalg.norm(cap[:, 1:], axis=1) <= 0.99]
    pts = np.concatenate([p, cap]).astype(np.float32)
    q = oracle.quantize(pts)
    flags, faces = oracle.hull(q)
    keep = []
    for o in range(0, len(q), 256):
        c = q[o:o + 256]
        try:
            f, fc = oracle.hull(c)
            flat = len(fc) == 0
        except RuntimeError:
            flat = True
        idx = np.arange(o, o + len(c))
        keep.append(idx if flat else idx[f])
    u = np.concatenate(keep)
    assert len(u) < len(q)
    f2, faces2 = oracle.hull(q[u])
    # same corners (as original indices) and the same facets
    assert np.array_equal(u[f2], np.nonzero(flags)[0])
    assert np.array_equal(u[faces2], faces)
    # a dropped flat chunk (the mistake the pass-through avoids) loses corners
    bad = np.concatenate([k for k, o in zip(keep, range(0, len(q), 256)) if o < len(p) - 256])
    fb, _ = oracle.hull(q[bad])
    assert not np.array_equal(bad[fb], np.nonzero(flags)[0])
