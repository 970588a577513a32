"""GPU parity tests of the exact hull against the CPU oracle.

Reference semantics: cuda/hull_dandc.cuh (quantized exact hull, lowest-index
duplicates, original points by index) and the reference's own hull tests
(tests/test_hull_mesh.py:90-237, tests/test_hull.py:494-743).
Bar: vertex index sets bit-exact on general-position clouds; FP64 volumes
within 1e-6 relative of the oracle's FP64 volume.
"""
import numpy as np
import pytest

import oracle
from cuacd_amd import datasets

pytestmark = pytest.mark.gpu

VOL_RTOL = 1e-6


def _index_sets_from_mesh(pts, hv):
    """Map output hull vertices back to input indices (exact float equality;
    the kernel copies input points bit for bit). Returns a sorted index array."""
    lut = {tuple(p.tolist()): i for i, p in reversed(list(enumerate(pts)))}
    return np.array(sorted(lut[tuple(v.tolist())] for v in hv))


def _check_mesh_closed(tris, nv):
    edges = {}
    for a, b, c in tris:
        for e in ((a, b), (b, c), (c, a)):
            edges[e] = edges.get(e, 0) + 1
    assert all(v == 1 for v in edges.values())
    assert all((b, a) in edges for (a, b) in edges)
    assert tris.min() >= 0 and tris.max() < nv


def test_cube_and_tetrahedron(ctx):
    cube = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
    tet = np.array([[1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]], np.float32)
    (cv, ct, cvol), (tv, tt, tvol) = ctx.batch_hull_dandc_mesh([cube, tet])
    assert len(cv) == 8 and len(ct) == 12
    assert len(tv) == 4 and len(tt) == 4
    assert abs(cvol - 1.0) < 1e-12 and abs(tvol - 8.0 / 3.0) < 1e-12
    _check_mesh_closed(ct, 8)
    vols, errs = ctx.batch_hull_volume([cube, tet])
    assert errs.tolist() == [0, 0]
    assert abs(vols[0] - 1.0) < 1e-6


def test_uniform_ball_parity(ctx):
    """1024 x 256 uniform-ball clouds (the batch_hull_volume workload)."""
    clouds = datasets.ball_clouds(1024, 256, seed=0)
    packed = np.concatenate(clouds, 0)
    off = np.arange(1025, dtype=np.int32) * 256
    oflags, onf, ovol, ogen = oracle.hull_batch(packed, off)
    res = ctx.batch_hull_dandc_mesh(clouds)
    vols, errs = ctx.batch_hull_volume_f64(clouds)
    assert (errs == 0).all()
    n_general = 0
    for i, (hv, ht, vol) in enumerate(res):
        _check_mesh_closed(ht, len(hv))
        assert vol > 0
        assert abs(vol - ovol[i]) <= VOL_RTOL * ovol[i], (i, vol, ovol[i])
        assert vols[i] == vol
        if ogen[i]:
            n_general += 1
            got = _index_sets_from_mesh(clouds[i], hv)
            want = np.nonzero(oflags[off[i]:off[i + 1]])[0]
            assert np.array_equal(got, want), i
            assert len(ht) == onf[i]
    assert n_general >= 1000, n_general


@pytest.mark.parametrize("n", [4, 5, 17, 64, 65, 200, 513, 1024, 2048])
def test_sizes_and_duplicates(ctx, n):
    rng = np.random.default_rng(n)
    p = rng.standard_normal((n, 3)).astype(np.float32)
    # duplicate a third of the points at higher indices
    p = np.concatenate([p, p[: n // 3]], 0)
    q = oracle.quantize(p)
    oflags, ofaces = oracle.hull(q)
    (hv, ht, vol), = ctx.batch_hull_dandc_mesh([p])
    assert abs(vol - oracle.mesh_volume(p, ofaces)) <= VOL_RTOL * abs(vol)
    if oracle.general_position(q[:n], ofaces):
        got = _index_sets_from_mesh(p, hv)
        assert np.array_equal(got, np.nonzero(oflags)[0])
        assert got.max() < n  # duplicates resolve to the lower index


def test_large_cloud_divide_and_conquer(ctx):
    """Clouds above one wave's reach are hulled chunk by chunk on the same grid."""
    rng = np.random.default_rng(5)
    clouds = [rng.standard_normal((n, 3)).astype(np.float32) for n in (3000, 20000)]
    res = ctx.batch_hull_dandc_mesh(clouds)
    for p, (hv, ht, vol) in zip(clouds, res):
        q = oracle.quantize(p)
        oflags, ofaces = oracle.hull(q)
        assert abs(vol - oracle.mesh_volume(p, ofaces)) <= VOL_RTOL * vol
        got = _index_sets_from_mesh(p, hv)
        assert np.array_equal(got, np.nonzero(oflags)[0]) or not oracle.general_position(q, ofaces)


@pytest.mark.parametrize("n", [1500, 6000, 16000])
def test_sphere_shell_block_tier(ctx, n):
    """Nearly every point is a hull vertex: the multi-wave tier with its workspace in HBM."""
    rng = np.random.default_rng(n)
    s = rng.standard_normal((n, 3))
    p = (s / np.linalg.norm(s, axis=1, keepdims=True) * (1 + 0.01 * rng.random((n, 1)))).astype(np.float32)
    (hv, ht, vol), = ctx.batch_hull_dandc_mesh([p])
    q = oracle.quantize(p)
    oflags, ofaces = oracle.hull(q)
    assert abs(vol - oracle.mesh_volume(p, ofaces)) <= VOL_RTOL * vol
    _check_mesh_closed(ht, len(hv))
    if oracle.general_position(q, ofaces):
        assert np.array_equal(_index_sets_from_mesh(p, hv), np.nonzero(oflags)[0])


def test_mixed_batch_all_tiers(ctx):
    """One batch mixing the one-wave, block and divide-and-conquer tiers."""
    rng = np.random.default_rng(9)
    clouds = [rng.standard_normal((n, 3)).astype(np.float32) for n in (10, 300, 1000, 1025, 5000, 40000, 77)]
    res = ctx.batch_hull_dandc_mesh(clouds)
    for p, (hv, ht, vol) in zip(clouds, res):
        _, ofaces = oracle.hull(oracle.quantize(p))
        assert abs(vol - oracle.mesh_volume(p, ofaces)) <= VOL_RTOL * vol


def test_fewer_than_four_points_and_flat(ctx):
    res = ctx.batch_hull_dandc_mesh([np.zeros((3, 3), np.float32),
                                     np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [.5, .5, 0]],
                                              np.float32)])
    for hv, ht, vol in res:
        assert len(ht) == 0 and vol == 0.0


def test_heap_stats_and_pool_usage_api(ctx):
    s = ctx.heap_stats()
    assert len(s) == 6
    assert ctx.pool_usage() >= 0


def test_error_injection_propagates(ctx):
    from cuacd_amd import _gpu
    _gpu.inject_error(1 << 16)
    cube = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    _, errs = ctx.batch_hull_volume([cube])
    assert errs[0] & (1 << 16)
    _, errs = ctx.batch_hull_volume([cube])
    assert errs[0] == 0


def _canonical_expect(p):
    q = oracle.quantize(p)
    flags, tris = oracle.hull(q)
    rank = -np.ones(len(p), np.int64)
    idx = np.nonzero(flags)[0]
    rank[idx] = np.arange(len(idx))
    return p[idx], rank[tris].astype(np.int32) if len(tris) else np.zeros((0, 3), np.int32), tris


@pytest.mark.parametrize("n,span", [(60, 2), (300, 3), (1000, 4), (1024, 6), (4000, 5), (12000, 8), (30000, 9)])
def test_canonical_mesh_equals_oracle_on_lattices(ctx, n, span):
    """Coplanar and collinear points everywhere (small integer lattices plus
    duplicates): the emitted mesh - corners, fan triangles, their order - is
    the oracle's canonical hull exactly, in every tier."""
    rng = np.random.default_rng(n + span)
    p = (rng.integers(0, span + 1, size=(n, 3)) * 0.25 - 0.3).astype(np.float32)
    (hv, ht, vol), = ctx.batch_hull_dandc_mesh([p])
    ev, et, ot = _canonical_expect(p)
    assert hv.tobytes() == ev.tobytes()
    assert np.array_equal(ht, et)
    # the volume's summation order is fixed too: bit for bit
    assert vol == oracle.hull_volume(p, ot)


def test_canonical_mesh_equals_oracle_on_cut_like_sets(ctx):
    """Half-ball point sets with a planar cap (the lookahead's typical hull
    input): the cap facet's points are coplanar on the grid."""
    rng = np.random.default_rng(11)
    outs = []
    for n in (200, 900, 3000):
        s = rng.standard_normal((n, 3))
        s /= np.linalg.norm(s, axis=1, keepdims=True)
        s = s[s[:, 2] > 0]
        ang = rng.random(n // 4) * 2 * np.pi
        cap = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1) * rng.random((n // 4, 1))
        outs.append(np.concatenate([s, cap]).astype(np.float32))
    res = ctx.batch_hull_dandc_mesh(outs)
    kres = ctx.batch_kdop_hull_mesh(outs)
    for p, (hv, ht, vol), (kv, kt, kvol) in zip(outs, res, kres):
        ev, et, ot = _canonical_expect(p)
        assert hv.tobytes() == ev.tobytes() and np.array_equal(ht, et)
        assert kv.tobytes() == ev.tobytes() and np.array_equal(kt, et)
        assert vol == oracle.hull_volume(p, ot)
