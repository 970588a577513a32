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
        assert vol[i] == oracle.mesh_volume(p, faces)
        assert gen[i] == int(oracle.general_position(q, faces))
    assert gen.sum() >= 24  # the uniform-ball protocol is overwhelmingly in general position


def test_planted_mistake_is_caught():
    """A wrong sign in the oracle's orientation would flip every face; the
    closed-form checks above must catch that: emulate by reversing faces."""
    p = _cube()
    _, f = oracle.hull(oracle.quantize(p))
    assert oracle.mesh_volume(p, f) > 0
    assert oracle.mesh_volume(p, f[:, ::-1].copy()) < 0
