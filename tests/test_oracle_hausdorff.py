"""CPU tests pinning the Hausdorff oracle (oracle/hausdorff_oracle.c).

Pins: closed-form point-triangle distances (face, edge, vertex regions), dense
brute-force sampling of the triangle, closed-form mesh cases from the
reference's own tests (tests/test_hausdorff.py:60-111), and the reference's
CoACD fixture (tests/data/hausdorff_coacd_ref.npz, criterion of
tests/test_hausdorff.py:126-155).
"""
import os

import numpy as np
import pytest

import oracle
from cuacd_amd import datasets

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

A, B, C = np.array([0., 0, 0]), np.array([1., 0, 0]), np.array([0., 1, 0])


@pytest.mark.parametrize("p,want", [((0.2, 0.2, 0.5), 0.25), ((-1, -1, 0), 2.0), ((0.5, -1, 0), 1.0),
                                    ((1, 1, 0), 0.5), ((2, 0, 1), 2.0), ((0.25, 0.25, 0), 0.0),
                                    ((0, 0.5, -3), 9.0)])
def test_point_triangle_closed_forms(p, want):
    assert abs(oracle.pt_tri_d2(np.array(p, float), A, B, C) - want) < 1e-12


def test_point_triangle_vs_dense_sampling():
    rng = np.random.default_rng(0)
    u, v = np.meshgrid(np.linspace(0, 1, 201), np.linspace(0, 1, 201))
    m = (u + v) <= 1
    u, v = u[m], v[m]
    for _ in range(50):
        a, b, c = rng.standard_normal((3, 3))
        p = rng.standard_normal(3) * 2
        pts = a + u[:, None] * (b - a) + v[:, None] * (c - a)
        dense = ((pts - p) ** 2).sum(1).min()
        d = oracle.pt_tri_d2(p, a, b, c)
        assert d <= dense + 1e-12
        assert dense - d < 0.02 * (1 + np.sqrt(dense))  # grid spacing bound


def test_mesh_closed_forms():
    outer = datasets.box_mesh((0, 0, 0), (2, 2, 2))
    inner = datasets.box_mesh((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))
    # corners of the outer box are sqrt(0.75) from the inner box
    assert abs(oracle.directed_hausdorff(outer[0], *inner) - np.sqrt(0.75)) < 1e-7
    # a box against itself
    assert oracle.directed_hausdorff(outer[0], *outer) == 0.0
    s = oracle.reference_samples(*outer)
    assert oracle.points_mesh_d2(s, *outer).max() < 1e-12  # samples lie on the surface
    # bidirectional value of concentric cubes is in the range the reference test asserts
    h = oracle.bidirectional_hausdorff_reference(*outer, *inner)
    assert 0.4 < h <= np.sqrt(0.75) + 1e-6


def test_reference_sampler_density():
    v, t = datasets.icosphere(3)
    v = v.astype(np.float32)
    counts, res = oracle.reference_sampler_counts(v, t)
    a, b, c = (v[t[:, k]].astype(np.float64) for k in range(3))
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    assert abs(res - 2000 * area) <= 2  # res = int(2000 * area)
    # per-triangle counts truncate: res - nt <= total <= res
    assert res - len(t) <= counts.sum() <= res


def test_coacd_fixture():
    d = np.load(os.path.join(GOLDEN, "hausdorff_coacd_ref.npz"), allow_pickle=False)
    for i in range(int(d["n_pairs"])):
        va, ta = d[f"verts_a_{i}"].astype(np.float32), d[f"tris_a_{i}"]
        vb, tb = d[f"verts_b_{i}"].astype(np.float32), d[f"tris_b_{i}"]
        ref = float(d[f"ref_{i}"])
        h = oracle.bidirectional_hausdorff_reference(va, ta, vb, tb)
        rel = abs(h - ref) / max(ref, 1e-6)
        assert rel < 0.15 or h <= 1.01 * ref, (i, h, ref)


def test_planted_mistake_is_caught():
    # a distance that forgot the edge regions (plane distance only) would
    # report 0 for a point beyond an edge in the triangle's plane
    assert oracle.pt_tri_d2(np.array([2.0, 2.0, 0.0]), A, B, C) > 1.0
