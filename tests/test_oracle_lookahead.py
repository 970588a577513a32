"""CPU checks of the lookahead oracle's pieces (leaf scoring, concave edges,
k-DOP axes, cost) and of the oracle search itself on small solids."""
import math

import numpy as np
import pytest

import oracle
import lookahead_oracle as LO
from cuacd_amd import datasets


def _norm(v):
    lo, hi = v.min(0), v.max(0)
    return ((v - (lo + hi) / 2) / ((hi - lo).max() / 2)).astype(np.float32)


def test_cost_closed_form():
    # rv of a volume gap g is the radius of the sphere of volume g
    g = 4.0 / 3.0 * math.pi * 0.2 ** 3
    assert abs(float(LO.cost_rv(1.0 + g, 1.0)) - 0.3 * 0.2) < 1e-7
    assert LO.cost_rv(1.0, 2.0) == 0.0


def test_kdop_axes():
    A = LO.kdop_axes().astype(np.float64)
    assert A.shape == (40, 3)
    assert np.allclose(np.linalg.norm(A, axis=1), 1.0, atol=1e-6)
    # no two axes parallel or antiparallel
    G = np.abs(A @ A.T) - np.eye(40)
    assert G.max() < 0.999
    # the 8 octant directions of the icosahedron's face normals are among them
    oct_ = np.abs(A) @ np.ones(3) / math.sqrt(3)
    assert (np.abs(oct_ - 1.0) < 1e-6).sum() == 4


@pytest.mark.parametrize("name,plane", [("x", (1, 0, 0, 0.1)), ("tilt", (0.3, -0.2, 0.9, 0.05)),
                                        ("diag", (0.7071, 0.7071, 0, -0.2))])
def test_score_cut_matches_capped_cut(name, plane):
    """Leaf scoring (no caps) gives the capped halves' volumes and the hull
    point sets whose hulls are the halves' hulls."""
    for v, t in (datasets.hollow_tube(), datasets.noisy_icospheres(1, 3, 0.15, seed=2)[0]):
        pl = np.float32(plane)
        pvol, nvol, npc, nnc, ppts, npts = LO.score_cut(v, t, pl)
        ref_p, ref_n = oracle.halfspace_volumes(v, t, pl)
        assert abs(pvol - ref_p) < 1e-6 * (ref_p + ref_n) and abs(nvol - ref_n) < 1e-6 * (ref_p + ref_n)
        st, pv, pt, nv, nt = oracle.plane_cut(v, t, pl)
        assert st == 0
        # same point sets as the capped halves' vertex sets
        assert sorted(map(tuple, ppts.tolist())) == sorted(map(tuple, pv.tolist()))
        assert sorted(map(tuple, npts.tolist())) == sorted(map(tuple, nv.tolist()))
        assert abs(LO.hull_volume(ppts)[0] - LO.hull_volume(pv)[0]) < 1e-12


def test_concave_edges_of_l_shape():
    v, t = datasets.l_shape_union()
    p = LO.Part(v, t, 3.0)
    planes = LO.concave_planes(p, 4, 0.005, float(np.float32(3.49)))
    # one concave edge (the notch, dihedral 3*pi/2): its two face planes and
    # the two offset bisector planes
    assert len(planes) == 4
    normals = sorted(tuple(np.round(q[:3], 4)) for q in planes)
    assert normals.count((0.7071, 0.7071, 0.0)) + normals.count((-0.7071, -0.7071, 0.0)) == 2
    box = datasets.box_mesh()
    assert LO.concave_planes(LO.Part(*box, 1.0), 4, 0.005, 3.49) == []


def test_oracle_search_l_shape_and_tube():
    for v, t in (datasets.l_shape_union(), datasets.hollow_tube()):
        v = _norm(v)
        hv, ht, _ = LO.hull_mesh(v)
        parts, iters = LO.lookahead(v, t, hv, ht, max_iters=20, width=12, width2=3, n_concave_edges=4)
        assert len(parts) >= 2
        total = abs(oracle.mesh_volume(v, t))
        vol = sum(abs(oracle.mesh_volume(p[0], p[1])) for p in parts)
        assert abs(vol - total) < 1e-5 * total
        for p in parts:
            assert oracle.edge_manifold_report(p[1]) == (0, 0)
            if iters < 20:
                assert max(LO.cost_rv(p[5], p[4]), p[6]) < np.float32(0.05)


def _quick_item(v, t, lc):
    v = np.asarray(v, np.float32)
    t = np.asarray(t, np.int32)
    h, _ = LO.hull_volume(v)
    return LO.Item(0, 3, [(LO.mesh_volume_abs(v, t), h, v, t, v.min(0), v.max(0))], list(lc))


def test_quick_level_most_balanced_axis():
    """The L-shape (box [0,2]x[0,1]x[0,1] plus box [0,1]x[1,2]x[0,1]), cut at
    its box midpoints: x and y split the volume 1 : 2 (cost 2/3), z splits it
    1.5 : 1.5 (cost 1/2), so the quick child is the z cut. Each half is an
    L-shaped slab whose hull exceeds it by 0.25, which fixes the level cost."""
    v, t = datasets.l_shape_union()
    extra = []
    out = LO.quick_levels([_quick_item(v, t, [np.float32(0.2)])], extra, 1, 3, np.float32(0.0125))
    assert len(out) == 1 and extra == []
    c = out[0]
    assert c.cut0 == 3 and not c.pruned and len(c.parts) == 2
    assert sorted((float(p[4][2]), float(p[5][2])) for p in c.parts) == [(0.0, 0.5), (0.5, 1.0)]
    assert all(abs(p[0] - 1.5) < 1e-12 for p in c.parts)
    rv = 0.3 * np.cbrt(3.0 / (4.0 * math.pi) * 0.25)
    assert len(c.lc) == 2 and c.lc[0] == np.float32(0.2) and abs(float(c.lc[1]) - rv) < 1e-6


def test_quick_level_tie_takes_first_axis_and_skips_thin_axes():
    """Flattened to 0.01 in z, the z axis spans less than 2 med + 1e-6 and is
    skipped; x and y tie at 2/3 and the first axis (x, midpoint 1) wins."""
    v, t = datasets.l_shape_union()
    v = np.asarray(v, np.float32).copy()
    v[:, 2] *= np.float32(0.01)
    out = LO.quick_levels([_quick_item(v, t, [np.float32(0.2)])], [], 1, 3, np.float32(0.0125))
    xs = sorted((float(p[4][0]), float(p[5][0]), round(p[0], 9)) for p in out[0].parts)
    assert xs == [(0.0, 1.0, 0.02), (1.0, 2.0, 0.01)]


def test_quick_level_small_part_is_extra_leaf():
    """A worst part spanning <= 2 med + 1e-6 along every axis is not cut: the
    item becomes an extra leaf with its level costs padded with zeros to the
    total number of levels, and with no child at all the level keeps the
    items as they are."""
    v, t = datasets.l_shape_union()
    v = np.asarray(v, np.float32) * np.float32(0.005)
    it = _quick_item(v, t, [np.float32(0.2)])
    extra = []
    out = LO.quick_levels([it], extra, 2, 4, np.float32(0.0125))
    assert out == [it]
    assert len(extra) == 1 and extra[0].lc == [np.float32(0.2), 0.0, 0.0, 0.0]


def test_oracle_search_with_quick_levels():
    """quick_depth 1 and 2 after two full levels: the decomposition still
    partitions the solid into closed parts below the threshold."""
    v, t = datasets.l_shape_union()
    v = _norm(v)
    hv, ht, _ = LO.hull_mesh(v)
    for q in (1, 2):
        parts, iters = LO.lookahead(v, t, hv, ht, max_iters=20, width=12, width2=3, n_concave_edges=4, quick_depth=q)
        assert iters < 20 and len(parts) >= 2
        total = abs(oracle.mesh_volume(v, t))
        assert abs(sum(abs(oracle.mesh_volume(p[0], p[1])) for p in parts) - total) < 1e-5 * total
        for p in parts:
            assert oracle.edge_manifold_report(p[1]) == (0, 0)
            assert max(LO.cost_rv(p[5], p[4]), p[6]) < np.float32(0.05)


def _bitrev8(x):
    return int(f"{x:08b}"[::-1], 2)


def test_merge_match_tie_rule():
    """The oracle's simulation of the reference's 256-thread matching equals
    its closed form: the cheapest pair under tau, ties to the smallest
    bit-reversed (index mod 256), then the smallest index; pairs sharing a
    part with a match drop out. Ties are dense here (four cost levels)."""
    rng = np.random.default_rng(5)
    levels = np.array([0.0, 0.01, 0.02, 1e20], np.float32)
    for P in (5, 30, 40, 60):
        n = P * (P - 1) // 2
        for _ in range(3):
            c = levels[rng.integers(0, 4, n)]
            got = LO.merge_match(c, 0.05)
            want, used = [], set()
            pairs = [(p1, p2) for p1 in range(1, P) for p2 in range(p1)]
            while True:
                ok = [i for i, (a, b) in enumerate(pairs) if c[i] < np.float32(0.05) and a not in used and b not in used]
                if not ok:
                    break
                i = min(ok, key=lambda i: (c[i], _bitrev8(i % 256), i))
                want.append(pairs[i])
                used.update(pairs[i])
            assert got == want
    # a tie the lowest-index rule would resolve the other way: index 254
    # (23, 1) and index 256 (23, 3) cost the same; thread 0 beats thread 254
    c = np.full(300, 1e20, np.float32)
    c[254] = c[256] = 0.0
    assert LO.merge_match(c, 0.05) == [(23, 3)]


def _part(lo, hi, cc=0):
    v, t = datasets.box_mesh(lo, hi)
    hv, ht, hvol = LO.hull_mesh(v)
    return LO.Part(v, t, LO.mesh_volume_abs(v, t), hvol, hv, ht, cc=cc)


def test_merge_pass_thin_slabs():
    """Two 1 x 1 x 0.02 slabs sharing the face x = 1 merge: their hull is the
    2 x 1 x 0.02 box (no volume gap) and the shared face lies within 0.01 of
    the hull's surface, under tau. The merged part is A's vertices and
    triangles, then B's, in A's slot (the higher index)."""
    a, b = _part((0, 0, 0), (1, 1, 0.02)), _part((1, 0, 0), (2, 1, 0.02))
    out = LO.merge_pass([a, b], 0.05)
    assert len(out) == 1
    m = out[0]
    assert np.array_equal(m.v, np.concatenate([b.v, a.v]))
    assert np.array_equal(m.t, np.concatenate([b.t, a.t + 8]))
    assert abs(m.mvol - 0.04) < 1e-9 and abs(m.hvol - 0.04) < 1e-9 and len(m.hv) == 8


def test_merge_pass_rejections():
    """No merge: slabs 0.03 apart (box gap >= tau / 2); unit cubes sharing a
    face (the shared face reaches 0.5 inside the merged hull, so the Hausdorff
    distance is 0.5); slabs of different connected components."""
    f = LO.merge_pass
    assert len(f([_part((0, 0, 0), (1, 1, 0.02)), _part((1.03, 0, 0), (2, 1, 0.02))], 0.05)) == 2
    assert len(f([_part((0, 0, 0), (1, 1, 1)), _part((1, 0, 0), (2, 1, 1))], 0.05)) == 2
    assert len(f([_part((0, 0, 0), (1, 1, 0.02), cc=1), _part((1, 0, 0), (2, 1, 0.02), cc=2)], 0.05)) == 2
    assert len(f([_part((0, 0, 0), (1, 1, 0.02), cc=1), _part((1, 0, 0), (2, 1, 0.02), cc=1)], 0.05)) == 1


def test_shared_plane_closed_form():
    """A unit cube's hull and a flat square on its face x = 1: that face's
    outward plane (1, 0, 0, 1) has the cube on its inner side and the square
    on it. Lifted 5e-4 off the face the square is still within the 1e-3
    tolerance; lifted 2e-3 it is not, and a square beside the cube (not on any
    face plane) finds nothing."""
    v, t = datasets.box_mesh()
    hv, ht, _ = LO.hull_mesh(v)
    sq = np.array([[1, 0.2, 0.2], [1, 0.8, 0.2], [1, 0.8, 0.8], [1, 0.2, 0.8]], np.float32)
    pl = LO.shared_plane(hv, ht, sq)
    assert pl is not None and np.allclose(pl, (1, 0, 0, 1), atol=1e-7)
    assert LO.shared_plane(hv, ht, sq + np.float32([5e-4, 0, 0])) is not None
    assert LO.shared_plane(hv, ht, sq + np.float32([2e-3, 0, 0])) is None
    assert LO.shared_plane(hv, ht, sq + np.float32([0.5, 0, 0])) is None
