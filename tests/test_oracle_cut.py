"""CPU checks of the plane-cut oracle (FP64 half-space volumes) and of the
voxel test solids, against closed forms and invariants.

The oracle clips triangles against the plane with the divergence origin on the
plane; a wrong clip (sign, interpolation, dropped piece) breaks one of the
closed forms or the conservation / symmetry invariants below.
"""
import numpy as np
import pytest

import oracle
from cuacd_amd import datasets


def _rot(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    return q * np.sign(np.diag(r))


@pytest.mark.parametrize("plane,expect", [
    ((0, 0, 1, -0.5), (0.5, 0.5)),
    ((1, 0, 0, -0.25), (0.75, 0.25)),
    ((0, -1, 0, 0.25), (0.25, 0.75)),
    ((1, 1, 1, -1.5), (0.5, 0.5)),            # through the centre, symmetric
    ((1, 1, 1, -0.5), (1 - 1 / 48, 1 / 48)),  # corner tetrahedron of volume (1/2)^3/6
    ((0, 0, 1, -2.0), (0.0, 1.0)),
])
def test_cube_closed_forms(plane, expect):
    v, t = datasets.cube_mesh()
    p, n = oracle.halfspace_volumes(v, t, plane)
    assert abs(p - expect[0]) < 1e-12 and abs(n - expect[1]) < 1e-12, (p, n)


def test_box_closed_forms():
    v, t = datasets.box_mesh((-1, -1.5, -2), (1, 1.5, 2))  # the 2x3x4 box
    for plane, pos in (((1, 0, 0, 0), 12.0), ((0, 1, 0, 0.5), 16.0), ((0, 0, 1, -0.3), 10.2)):
        p, n = oracle.halfspace_volumes(v, t, plane)
        assert abs(p - pos) < 1e-12 and abs(n - (24 - pos)) < 1e-12


def test_solids_closed():
    for f, vol in ((datasets.hollow_tube, 6.0), (datasets.two_holes, 15.0), (datasets.l_shape_union, 3.0)):
        v, t = f()
        assert oracle.edge_manifold_report(t) == (0, 0)
        p, n = oracle.halfspace_volumes(v, t, (0.3, -0.2, 0.9, 0.05))
        assert abs(p + n - vol) < 1e-9


def test_solid_closed_forms():
    v, t = datasets.l_shape_union()
    assert np.allclose(oracle.halfspace_volumes(v, t, (0, 1, 0, -0.5)), (2.0, 1.0))
    v, t = datasets.hollow_tube()
    assert np.allclose(oracle.halfspace_volumes(v, t, (1, 0, 0, -0.75)), (1.0, 5.0))  # 0.25 x 2 x 2 slab
    v, t = datasets.two_holes()
    # x >= 1: 1x2x2 box minus half of the hole at x = 1 (0.25 x 0.5 x 2)
    assert np.allclose(oracle.halfspace_volumes(v, t, (1, 0, 0, -1)), (4 - 0.25, 11.25))


def test_conservation_and_symmetry_random():
    meshes = datasets.noisy_icospheres(4, 3, 0.15, seed=9)
    rng = np.random.default_rng(1)
    for v, t in meshes:
        total = oracle.mesh_volume(v, t)
        for _ in range(8):
            nrm = rng.standard_normal(3)
            nrm /= np.linalg.norm(nrm)
            plane = (*nrm, float(rng.uniform(-0.5, 0.5)))
            p, n = oracle.halfspace_volumes(v, t, plane)
            assert abs(p + n - total) < 1e-9 * total
            p2, n2 = oracle.halfspace_volumes(v, t, tuple(-x for x in plane))
            assert abs(p - n2) < 1e-9 and abs(n - p2) < 1e-9
            # scaling the plane equation changes nothing
            p3, _ = oracle.halfspace_volumes(v, t, tuple(3.0 * x for x in plane))
            assert abs(p - p3) < 1e-9


def test_rotation_invariance():
    v, t = datasets.two_holes()
    R = _rot(3)
    vr = (v.astype(np.float64) @ R.T)
    p, n = oracle.halfspace_volumes(v, t, (0, 0, 1, -0.1))
    nr = R @ np.array([0, 0, 1.0])
    pr, nrv = oracle.halfspace_volumes(vr.astype(np.float32), t, (*nr, -0.1))
    assert abs(p - pr) < 1e-5 and abs(n - nrv) < 1e-5


def test_manifold_report_detects_defects():
    v, t = datasets.cube_mesh()
    assert oracle.edge_manifold_report(t[:-1]) == (3, 0)         # open
    bad = t.copy()
    bad[0] = bad[0][::-1]
    o, d = oracle.edge_manifold_report(bad)
    assert d > 0                                               # flipped triangle


# ---- the serial plane cut with caps ------------------------------------------------

def _signed_volume(v, t):
    v = v.astype(np.float64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum()) / 6.0


def _cut_cases():
    cube = datasets.box_mesh((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    box = datasets.box_mesh((-1, -1.5, -2), (1, 1.5, 2))
    occ = np.ones((5, 5, 5), bool)
    occ[1:4, 1:4, 1:4] = False
    occ[2, 2, 2] = True
    nested = datasets.voxel_solid(occ, (0, 0, 0), 1.0)
    sph = datasets.icosphere(3)
    return [
        ("cube_z", cube, (0, 0, 1, 0)), ("cube_off", cube, (0, 0, 1, -0.25)),
        ("cube_diag", cube, (0.7071, 0.7071, 0, 0)), ("cube_vertex", cube, (1, 1, 1, -1.5)),
        ("box_y", box, (0, 1, 0, 0.5)), ("box_z", box, (0, 0, 1, -0.3)),
        ("sphere", (sph[0].astype(np.float32), sph[1]), (0, 0, 1, 0)),
        ("sphere_tilt", (sph[0].astype(np.float32), sph[1]), (0.3, -0.5, 0.8, 0.1)),
        ("tube_ring", datasets.hollow_tube(), (0, 0, 1, -0.1)),
        ("tube_lattice", datasets.hollow_tube(), (0, 0, 1, 0)),
        ("tube_two_caps", datasets.hollow_tube(), (1, 0, 0, 0)),
        ("two_holes", datasets.two_holes(), (0.05, -0.1, 1, -0.3)),
        ("two_holes_lattice", datasets.two_holes(), (0, 0, 1, 0)),
        ("l_notch", datasets.l_shape_union(), (0, 0, 1, -0.5)),
        ("l_diag", datasets.l_shape_union(), (1, 1, 0, -1.7)),
        ("nested", nested, (0.1, 0.05, 1, -2.4)),
        ("nested_lattice", nested, (0, 0, 1, -2.5)),
    ]


@pytest.mark.parametrize("name,mesh,plane", _cut_cases(), ids=[c[0] for c in _cut_cases()])
def test_oracle_plane_cut_closed_and_exact_volume(name, mesh, plane):
    v, t = mesh
    st, pv, pt, nv, nt = oracle.plane_cut(v, t, plane)
    assert st in (0, 2), st
    total = abs(oracle.mesh_volume(v, t))
    ref_p, ref_n = oracle.halfspace_volumes(v, t, np.float32(plane))
    for hv, ht, ref in ((pv, pt, ref_p), (nv, nt, ref_n)):
        if len(ht) == 0:
            assert abs(ref) < 1e-6 * total
            continue
        assert oracle.edge_manifold_report(ht) == (0, 0), name
        assert abs(_signed_volume(hv, ht) - ref) < 1e-5 * total, (name, _signed_volume(hv, ht), ref)


def test_oracle_plane_cut_random_noisy_spheres():
    meshes = datasets.noisy_icospheres(6, 3, 0.15, seed=21)
    rng = np.random.default_rng(2)
    for v, t in meshes:
        for _ in range(6):
            n = rng.standard_normal(3)
            n /= np.linalg.norm(n)
            plane = np.float32([*n, rng.uniform(-0.7, 0.7)])
            st, pv, pt, nv, nt = oracle.plane_cut(v, t, plane)
            assert st in (0, 2)
            ref_p, ref_n = oracle.halfspace_volumes(v, t, plane)
            for hv, ht, ref in ((pv, pt, ref_p), (nv, nt, ref_n)):
                if len(ht):
                    assert oracle.edge_manifold_report(ht) == (0, 0)
                    assert abs(_signed_volume(hv, ht) - ref) < 1e-5 * (ref_p + ref_n)


def test_oracle_plane_cut_sdf_meshes():
    from cuacd_amd import sdf
    for v, t in sdf.sdf_meshes(2, seed=3, res=24):
        for plane in ((1, 0, 0, 0.013), (0.2, 0.9, -0.3, -0.05)):
            plane = np.float32(plane)
            st, pv, pt, nv, nt = oracle.plane_cut(v, t, plane)
            assert st in (0, 2)
            ref_p, ref_n = oracle.halfspace_volumes(v, t, plane)
            for hv, ht, ref in ((pv, pt, ref_p), (nv, nt, ref_n)):
                if len(ht):
                    assert oracle.edge_manifold_report(ht) == (0, 0)
                    assert abs(_signed_volume(hv, ht) - ref) < 1e-5 * (ref_p + ref_n)
