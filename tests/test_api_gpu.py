"""GPU tests of check_mesh, batch_kdop_hull_mesh and the connected-components
entry (reference cuda/mesh_audit.cuh, cuda/kdop_hull.cuh, tests/test_decompose_components.py)."""
import numpy as np
import pytest

import oracle
from cuacd_amd import _gpu, datasets, sdf

pytestmark = pytest.mark.gpu


def _cube(center=(0, 0, 0), size=1.0):
    v, t = datasets.box_mesh(tuple(np.subtract(center, size / 2)), tuple(np.add(center, size / 2)))
    return v, t


def _merge(ms):
    vs, ts, off = [], [], 0
    for v, t in ms:
        vs.append(v)
        ts.append(t + off)
        off += len(v)
    return np.concatenate(vs).astype(np.float32), np.concatenate(ts).astype(np.int32)


def test_check_mesh_verdicts(ctx):
    v, t = _cube()
    r = ctx.check_mesh(v, t)
    assert r["flags"] == 0 and r["watertight"] and r["manifold"] and r["oriented"] and not r["needs_remesh"]
    r = ctx.check_mesh(v, t[:-1])
    assert not r["watertight"] and r["needs_remesh"]
    bad = t.copy()
    bad[0] = bad[0][::-1]
    r = ctx.check_mesh(v, bad)
    assert not r["oriented"]
    r = ctx.check_mesh(v, np.vstack([t, [[0, 1, 99]]]))
    assert r["bad_index"]
    r = ctx.check_mesh(v, np.vstack([t, [[0, 0, 1]]]))
    assert r["degenerate"]
    # three triangles on one edge
    r = ctx.check_mesh(v, np.vstack([t, [[0, 1, 6]]]))
    assert not r["manifold"]
    for v, t in sdf.sdf_meshes(3, seed=2, res=24):
        assert ctx.check_mesh(v, t)["flags"] == 0
        assert oracle.edge_manifold_report(t) == (0, 0)


@pytest.mark.parametrize("n", [300, 1024, 1025, 3000, 12000])
def test_kdop_hull_matches_exact_oracle(ctx, n):
    rng = np.random.default_rng(n)
    clouds = []
    for k in range(3):
        p = rng.standard_normal((n, 3))
        if k == 1:
            p /= np.linalg.norm(p, axis=1, keepdims=True)  # sphere shell: every point on the hull
        clouds.append(p.astype(np.float32))
    res = ctx.batch_kdop_hull_mesh(clouds)
    for p, (hv, ht, vol) in zip(clouds, res):
        flags, faces = oracle.hull(oracle.quantize(p))
        ref = abs(oracle.mesh_volume(p, faces))
        assert abs(vol - ref) <= 1e-6 * ref, (vol, ref)
        assert sorted(map(tuple, hv.tolist())) == sorted(map(tuple, p[flags].tolist()))
        assert oracle.edge_manifold_report(ht) == (0, 0)


def _dc(verts, tris, max_comp=64):
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    nv, nt = len(verts), len(tris)
    mvp, mtp = nv + 16, nt + 16
    ov = np.zeros((max_comp * mvp, 3), np.float32)
    ot = np.zeros((max_comp * mtp, 3), np.int32)
    onv = np.zeros(max_comp, np.int32)
    ont = np.zeros(max_comp, np.int32)
    n = _gpu.test_postprocess_dc(verts.ctypes.data, nv, tris.ctypes.data, nt, max_comp, mvp, mtp,
                                 ov.ctypes.data, ot.ctypes.data, onv.ctypes.data, ont.ctypes.data)
    return [(ov[c * mvp:c * mvp + onv[c]].copy(), ot[c * mtp:c * mtp + ont[c]].copy()) for c in range(n)]


def _sv(v, t):
    return float(np.sum(v[t[:, 0]] * np.cross(v[t[:, 1]], v[t[:, 2]]))) / 6.0


def test_components_reference_cases(ctx):
    c = _dc(*_cube())
    assert len(c) == 1 and len(c[0][0]) == 8 and len(c[0][1]) == 12
    c = _dc(*_merge([_cube((0, 0, 0)), _cube((5, 0, 0))]))
    assert sorted(len(x[0]) for x in c) == [8, 8] and sorted(len(x[1]) for x in c) == [12, 12]
    tet = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32),
           np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.int32))
    c = _dc(*_merge([(tet[0] + o, tet[1]) for o in ((0, 0, 0), (10, 0, 0), (0, 10, 0))]))
    assert len(c) == 3 and all(len(x[0]) == 4 and len(x[1]) == 4 for x in c)
    # two cubes sharing a face stay one component
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1],
                  [2, 0, 0], [2, 1, 0], [2, 0, 1], [2, 1, 1]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [2, 3, 7], [2, 7, 6],
                  [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5], [1, 9, 8], [1, 2, 9], [5, 10, 11], [5, 11, 6],
                  [1, 8, 10], [1, 10, 5], [9, 2, 6], [9, 6, 11], [1, 5, 6], [1, 6, 2], [8, 9, 11], [8, 11, 10]],
                 np.int32)
    assert len(_dc(v, t)) == 1
    v, t = _merge([_cube((0, 0, 0), 2.0), _cube((10, 0, 0), 3.0)])
    c = _dc(v, t)
    assert abs(sum(abs(_sv(*x)) for x in c) - abs(_sv(v, t))) < 1e-3
    c = _dc(*_merge([_cube((3 * i, 0, 0)) for i in range(33)]))
    assert len(c) == 33
    assert all(abs(_sv(*x) - 1.0) < 1e-3 for x in c)


def test_components_drop_inner_shell(ctx):
    """A hollow box (outer shell plus an inward-facing inner shell) keeps only
    the outer one when the pieces' orientations mix."""
    outer = _cube((0, 0, 0), 4.0)
    inner = _cube((0, 0, 0), 1.0)
    inner = (inner[0], inner[1][:, ::-1].copy())
    c = _dc(*_merge([outer, inner]))
    assert len(c) == 1 and abs(_sv(*c[0]) - 64.0) < 1e-3
