"""GPU tests of lookahead_decompose (reference tests/test_lookahead.py:124-187
and the search semantics of csrc/lookahead.c / cuda/la_*.cu).

Every run checks what any decomposition must satisfy: the parts are closed
meshes, their volumes add up to the input's, and (unless the iteration cap
was hit) every part's cost max(0.3 rv, H) is below the threshold.
"""
import numpy as np
import pytest

import oracle
import os, sys
sys.path.insert(0, os.path.dirname(__file__))
from cuacd_amd import datasets, sdf

pytestmark = pytest.mark.gpu


def _normalize(v):
    lo, hi = v.min(0), v.max(0)
    c = (lo + hi) / 2
    s = (hi - lo).max() / 2
    return ((v - c) / s).astype(np.float32)


def _check_parts(parts, v, t, info=None, tau=0.05, rel=1e-4):
    total = abs(oracle.mesh_volume(v, t))
    vol = 0.0
    for pv, pt, hv, ht in parts:
        assert len(pt) >= 4 and len(ht) >= 4
        assert oracle.edge_manifold_report(pt) == (0, 0)
        vol += abs(oracle.mesh_volume(pv, pt))
        # the hull encloses the part up to its quantization grid: the hull is
        # exact on the integer grid (per-axis step span / 10216, truncation),
        # so a vertex left out lies within one grid diagonal of the quantized
        # hull, which lies within one grid diagonal of the float hull - the
        # part pokes out by at most a shell of that thickness over the hull
        step = float((pv.max(0) - pv.min(0)).max()) / 10216.0
        e = (hv[ht[:, 1]] - hv[ht[:, 0]]).astype(np.float64)
        f = (hv[ht[:, 2]] - hv[ht[:, 0]]).astype(np.float64)
        area = 0.5 * np.linalg.norm(np.cross(e, f), axis=1).sum()
        slack = 2.0 * np.sqrt(3.0) * step * area
        assert abs(oracle.mesh_volume(hv, ht)) >= abs(oracle.mesh_volume(pv, pt)) - slack
    assert abs(vol - total) <= rel * total, (vol, total)
    if info is not None:
        rv = 0.3 * np.cbrt(3 / (4 * np.pi) * np.maximum(info["hull_vol"] - info["mesh_vol"], 0))
        cost = np.maximum(rv, info["hausdorff"])
        return cost
    return None


def test_cube_is_one_part(ctx):
    v, t = datasets.cube_mesh()
    parts = ctx.lookahead_decompose(_normalize(v), t, max_iters=5, threshold=0.05)
    assert len(parts) == 1


def test_lshape(ctx):
    v, t = datasets.l_shape_union()
    v = _normalize(v)
    parts = ctx.lookahead_decompose(v, t, max_iters=50, threshold=0.05)
    assert 2 <= len(parts) <= 10
    _check_parts(parts, v, t)


def test_batch_matches_single_and_converges(ctx):
    meshes = [(_normalize(v), t) for v, t in sdf.sdf_meshes(3, seed=7, res=24)]
    meshes.append((_normalize(datasets.l_shape_union()[0]), datasets.l_shape_union()[1]))
    kw = dict(max_iters=100, width=30, width2=5, threshold=0.05, depth=2, n_concave_edges=16)
    batch, info = ctx.batch_lookahead_decompose(meshes, return_info=True, **kw)
    for m, (v, t) in enumerate(meshes):
        cost = _check_parts(batch[m], v, t, info[m])
        if info[m]["iters"] < kw["max_iters"]:
            assert (cost < kw["threshold"]).all(), (m, cost.max())
        single = ctx.lookahead_decompose(v, t, **kw)
        assert len(single) == len(batch[m])
        for (a, b, _, _), (c, d, _, _) in zip(single, batch[m]):
            assert a.tobytes() == c.tobytes() and np.array_equal(b, d)


def test_heap_balanced(ctx):
    before = ctx.heap_stats()
    v, t = datasets.hollow_tube()
    ctx.lookahead_decompose(_normalize(v), t, max_iters=8, width=12, width2=3, n_concave_edges=4)
    after = ctx.heap_stats()
    assert after[0] == before[0] and after[3] == before[3], (before, after)


def _rotated(v, seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return (v.astype(np.float64) @ q.T).astype(np.float32)


def test_parity_with_serial_oracle(ctx):
    """The GPU search makes the oracle's decisions: same iterations, same
    parts (vertex bits and triangles) on general-position inputs."""
    import lookahead_oracle as LO
    cases = [(_rotated(datasets.l_shape_union()[0], 1), datasets.l_shape_union()[1]),
             (_rotated(datasets.hollow_tube()[0], 2), datasets.hollow_tube()[1])]
    cases += sdf.sdf_meshes(2, seed=11, res=16)
    cases.append(datasets.noisy_icospheres(1, 4, 0.1, seed=3)[0])
    kw = dict(max_iters=12, width=12, width2=3, threshold=0.05, depth=2, n_concave_edges=4, max_n_cutting=4)
    for ci, (v, t) in enumerate(cases):
        v = _normalize(v)
        if ci == len(cases) - 1:
            kw = dict(kw, max_iters=3)
        hv, ht = ctx._input_hull(v)
        gpu, info = ctx.batch_lookahead_decompose([(v, t)], return_info=True, **kw)
        ref, iters = LO.lookahead(v, t, hv, ht, **kw)
        assert info[0]["iters"] == iters, (ci, info[0]["iters"], iters)
        assert len(gpu[0]) == len(ref), (ci, len(gpu[0]), len(ref))
        key = lambda p: (len(p[1]), p[0].tobytes())
        g = sorted(gpu[0], key=key)
        r = sorted(ref, key=key)
        for (gv, gt, ghv, ght), rp in zip(g, r):
            assert gv.tobytes() == rp[0].tobytes() and np.array_equal(gt, rp[1]), ci
            # Both hulls are exact on the quantization grid, but where grid
            # points are coplanar the two algorithms may keep different
            # face-interior points as vertices; the float volumes then differ
            # by the points' distance from the grid plane (~1e-8 relative).
            assert abs(abs(oracle.mesh_volume(ghv, ght)) - rp[5]) <= 1e-6 * max(rp[5], 1e-12) + 1e-12, ci


def test_merge_hulls_pass(ctx):
    """merge_hulls=True only merges pairs whose merged cost stays under the
    threshold: fewer (or as many) parts, same total volume, every merged hull
    encloses its part."""
    meshes = [(_normalize(v), t) for v, t in sdf.sdf_meshes(2, seed=21, res=20)]
    kw = dict(max_iters=40, width=30, width2=5, threshold=0.05, depth=2, n_concave_edges=16)
    plain = ctx.batch_lookahead_decompose(meshes, **kw)
    merged, info = ctx.batch_lookahead_decompose(meshes, merge_hulls=True, return_info=True, **kw)
    for m, (v, t) in enumerate(meshes):
        assert len(merged[m]) <= len(plain[m])
        total = abs(oracle.mesh_volume(v, t))
        vol = sum(abs(oracle.mesh_volume(p[0], p[1])) for p in merged[m])
        assert abs(vol - total) <= 1e-4 * total
        rv = 0.3 * np.cbrt(3 / (4 * np.pi) * np.maximum(info[m]["hull_vol"] - info[m]["mesh_vol"], 0))
        if info[m]["iters"] < kw["max_iters"]:
            assert (rv < kw["threshold"]).all()


def test_parity_with_serial_oracle_bench_parameters(ctx):
    """The benchmark's search parameters (width/width2 30/5, 16 concave edges,
    16 parts cut per iteration) on two SDF solids, three iterations: the GPU
    search and the serial oracle make the same decisions."""
    import lookahead_oracle as LO
    kw = dict(max_iters=3, width=30, width2=5, threshold=0.05, depth=2, n_concave_edges=16, max_n_cutting=16)
    for ci, (v, t) in enumerate(sdf.sdf_meshes(2, seed=5, res=20)):
        v = _normalize(v)
        hv, ht = ctx._input_hull(v)
        gpu, info = ctx.batch_lookahead_decompose([(v, t)], return_info=True, **kw)
        ref, iters = LO.lookahead(v, t, hv, ht, **kw)
        assert info[0]["iters"] == iters, (ci, info[0]["iters"], iters)
        key = lambda p: (len(p[1]), p[0].tobytes())
        g = sorted(gpu[0], key=key)
        r = sorted(ref, key=key)
        assert len(g) == len(r), (ci, len(g), len(r))
        for (gv, gt, ghv, ght), rp in zip(g, r):
            assert gv.tobytes() == rp[0].tobytes() and np.array_equal(gt, rp[1]), ci
            assert ghv.tobytes() == rp[2].tobytes() and np.array_equal(ght, rp[3]), ci


@pytest.mark.parametrize("quick,depth", [(1, 2), (2, 2), (1, 1), (2, 1)])
def test_parity_with_serial_oracle_quick_levels(ctx, quick, depth):
    """Quick levels after the full ones (csrc/lookahead.c:656-727,
    cuda/la_refine.cu:52-351): the batched GPU search over four solids makes
    the serial oracle's decisions for each, part for part, and frees every
    half it built; the single-mesh entry agrees with the batch."""
    import lookahead_oracle as LO
    cases = [(_rotated(datasets.l_shape_union()[0], 1), datasets.l_shape_union()[1]),
             (_rotated(datasets.hollow_tube()[0], 2), datasets.hollow_tube()[1])]
    cases += sdf.sdf_meshes(2, seed=11, res=16)
    cases = [(_normalize(v), t) for v, t in cases]
    kw = dict(max_iters=12, width=12, width2=3, threshold=0.05, depth=depth, quick_depth=quick, n_concave_edges=4,
              max_n_cutting=4)
    print(f"quick levels: quick_depth {quick} depth {depth}", flush=True)
    before = ctx.heap_stats()
    gpu, info = ctx.batch_lookahead_decompose(cases, return_info=True, **kw)
    after = ctx.heap_stats()
    assert after[0] == before[0] and after[3] == before[3], (before, after)  # every quick half freed
    for ci, (v, t) in enumerate(cases):
        hv, ht = ctx._input_hull(v)
        ref, iters = LO.lookahead(v, t, hv, ht, **kw)
        assert info[ci]["iters"] == iters, (ci, info[ci]["iters"], iters)
        assert len(gpu[ci]) == len(ref), (ci, len(gpu[ci]), len(ref))
        key = lambda p: (len(p[1]), p[0].tobytes())
        g = sorted(gpu[ci], key=key)
        r = sorted(ref, key=key)
        for (gv, gt, ghv, ght), rp in zip(g, r):
            assert gv.tobytes() == rp[0].tobytes() and np.array_equal(gt, rp[1]), ci
            assert abs(abs(oracle.mesh_volume(ghv, ght)) - rp[5]) <= 1e-6 * max(rp[5], 1e-12) + 1e-12, ci
        if info[ci]["iters"] < kw["max_iters"]:
            cost = _check_parts(gpu[ci], v, t, info[ci])
            assert (cost < kw["threshold"]).all(), (ci, cost.max())
    single = ctx.lookahead_decompose(*cases[2], **kw)
    assert len(single) == len(gpu[2])
    for (a, b, _, _), (c, d, _, _) in zip(single, gpu[2]):
        assert a.tobytes() == c.tobytes() and np.array_equal(b, d)


def test_parity_with_serial_oracle_merge(ctx):
    """merge_hulls=True (csrc/lookahead.c:988-1067, cuda/postprocess_merge.cu):
    on four solids where the pass merges 1-9 pairs, the batched GPU search
    and the serial oracle end with the same parts, merged ones included
    (vertex bits, triangles, hull volume), and every chunk is freed."""
    import lookahead_oracle as LO
    cases = [(_rotated(datasets.hollow_tube()[0], 2), datasets.hollow_tube()[1])]
    cases += sdf.sdf_meshes(4, seed=21, res=16)[:3]
    cases = [(_normalize(v), t) for v, t in cases]
    kw = dict(max_iters=12, width=12, width2=3, threshold=0.05, depth=2, n_concave_edges=4, max_n_cutting=4,
              merge_hulls=True)
    before = ctx.heap_stats()
    gpu, info = ctx.batch_lookahead_decompose(cases, return_info=True, **kw)
    after = ctx.heap_stats()
    assert after[0] == before[0] and after[3] == before[3], (before, after)
    plain = ctx.batch_lookahead_decompose(cases, **dict(kw, merge_hulls=False))
    merged_any = False
    for ci, (v, t) in enumerate(cases):
        hv, ht = ctx._input_hull(v)
        ref, iters = LO.lookahead(v, t, hv, ht, **kw)
        assert info[ci]["iters"] == iters, (ci, info[ci]["iters"], iters)
        assert len(gpu[ci]) == len(ref), (ci, len(gpu[ci]), len(ref))
        merged_any |= len(gpu[ci]) < len(plain[ci])
        key = lambda p: (len(p[1]), p[0].tobytes())
        g = sorted([p + (info[ci]["hull_vol"][k],) for k, p in enumerate(gpu[ci])], key=key)
        r = sorted(ref, key=key)
        for k, ((gv, gt, ghv, ght, ghvol), rp) in enumerate(zip(g, r)):
            assert gv.tobytes() == rp[0].tobytes() and np.array_equal(gt, rp[1]), (ci, k)
            assert abs(float(ghvol) - rp[5]) <= 1e-6 * max(rp[5], 1e-12) + 1e-12, (ci, k)
            assert abs(abs(oracle.mesh_volume(ghv, ght)) - rp[5]) <= 1e-6 * max(rp[5], 1e-12) + 1e-12, (ci, k)
    assert merged_any


@pytest.mark.parametrize("dc_final,dc_iter", [(True, True), (False, False), (True, False)])
def test_parity_with_serial_oracle_component_options(ctx, dc_final, dc_iter):
    """decompose_components (a components pass after the loop) and
    no_decompose_components_per_iter (no components pass inside it),
    reference csrc/lookahead.c:336-345, 909-931, 969-985: the batched GPU
    search and the serial oracle end with the same parts on three solids."""
    import lookahead_oracle as LO
    cases = [(_rotated(datasets.hollow_tube()[0], 2), datasets.hollow_tube()[1])]
    cases += sdf.sdf_meshes(2, seed=31, res=16)
    cases = [(_normalize(v), t) for v, t in cases]
    kw = dict(max_iters=10, width=12, width2=3, threshold=0.05, depth=2, n_concave_edges=4, max_n_cutting=4)
    print(f"component options: decompose_components {dc_final} per-iteration {dc_iter}", flush=True)
    gpu, info = ctx.batch_lookahead_decompose(cases, return_info=True, decompose_components=dc_final,
                                              no_decompose_components_per_iter=not dc_iter, **kw)
    for ci, (v, t) in enumerate(cases):
        hv, ht = ctx._input_hull(v)
        ref, iters = LO.lookahead(v, t, hv, ht, decompose_components=dc_final, dc_per_iter=dc_iter, **kw)
        assert info[ci]["iters"] == iters, (ci, info[ci]["iters"], iters)
        assert len(gpu[ci]) == len(ref), (ci, len(gpu[ci]), len(ref))
        key = lambda p: (len(p[1]), p[0].tobytes())
        g = sorted(gpu[ci], key=key)
        r = sorted(ref, key=key)
        for (gv, gt, ghv, ght), rp in zip(g, r):
            assert gv.tobytes() == rp[0].tobytes() and np.array_equal(gt, rp[1]), ci
            assert abs(abs(oracle.mesh_volume(ghv, ght)) - rp[5]) <= 1e-6 * max(rp[5], 1e-12) + 1e-12, ci


def test_quick_level_error_paths(ctx):
    """quick_depth < 0 is rejected; depth + quick_depth beyond the item's
    part capacity (8 parts: the reference's KERR_LA_OVERFLOW,
    cuda/la_refine.cu:267-274) is an error, not a silent truncation; the
    heap stays balanced either way."""
    v, t = datasets.l_shape_union()
    v = _normalize(_rotated(v, 1))
    before = ctx.heap_stats()
    with pytest.raises(Exception):
        ctx.batch_lookahead_decompose([(v, t)], max_iters=3, width=12, width2=3, n_concave_edges=4, quick_depth=-1)
    with pytest.raises(Exception, match="capacity"):
        ctx.batch_lookahead_decompose([(v, t)], max_iters=3, width=12, width2=3, n_concave_edges=4, depth=2,
                                      quick_depth=6)
    after = ctx.heap_stats()
    assert after[0] == before[0] and after[3] == before[3], (before, after)


@pytest.mark.parametrize("name,kw", [
    ("chevron_d1", dict(width=9, width2=5, depth=1, n_concave_edges=0)),
    ("chevron_d2", dict(width=9, width2=5, depth=2, n_concave_edges=0)),
    ("lshape_edges_d2", dict(width=12, width2=5, depth=2, n_concave_edges=4)),
    ("comb_reservoir", dict(width=12, width2=3, depth=2, n_concave_edges=2, max_n_cutting=4)),
    ("ushape_defaults", dict()),
])
def test_parity_on_closed_form_solids(ctx, name, kw):
    """Solids whose first decisions the oracle's closed-form tests fix (the
    chevron's convex split, the L-shape's concave-plane cut and its mean path
    cost, the comb's reservoir replacing sampled edges): the GPU search makes
    the same decisions, part for part, bit for bit."""
    import lookahead_oracle as LO
    solid = {"chevron": datasets.chevron_prism, "lshape": datasets.l_shape_union,
             "comb": lambda: datasets.comb_solid(4), "ushape": datasets.u_shape}[name.split("_")[0]]
    v, t = solid()
    kw = dict(dict(max_iters=20, threshold=0.05), **kw)
    hv, ht = ctx._input_hull(v)
    gpu, info = ctx.batch_lookahead_decompose([(v, t)], return_info=True, **kw)
    ref, iters = LO.lookahead(v, t, hv, ht, **kw)
    assert info[0]["iters"] == iters
    key = lambda p: (len(p[1]), p[0].tobytes())
    g = sorted(gpu[0], key=key)
    r = sorted(ref, key=key)
    assert len(g) == len(r)
    for (gv, gt, ghv, ght), rp in zip(g, r):
        assert gv.tobytes() == rp[0].tobytes() and np.array_equal(gt, rp[1])
    if name.startswith("chevron"):
        # the convex split: two halves of volume 1.5 in one iteration
        assert iters == 1 and len(g) == 2
        assert np.allclose(sorted(info[0]["mesh_vol"]), [1.5, 1.5], atol=1e-6)


def test_cut_paths_give_the_same_search(ctx):
    """The level-0 cuts on the LDS-resident path and on the heap-scratch path
    (knob cut_lds_verts) drive the search to the same parts and hulls, bit
    for bit."""
    from cuacd_amd import sdf
    meshes = sdf.sdf_meshes(3, seed=12, res=24)
    kw = dict(max_iters=4, width=30, width2=5, threshold=0.05, depth=2, max_n_cutting=16, n_concave_edges=16)
    old = ctx.get_tuning("cut_lds_verts")
    out = []
    try:
        for cap in (98304, 0):
            ctx.set_tuning("cut_lds_verts", cap)
            out.append(ctx.batch_lookahead_decompose(meshes, **kw))
    finally:
        ctx.set_tuning("cut_lds_verts", old)
    for pa, pb in zip(*out):
        assert len(pa) == len(pb)
        for x, y in zip(pa, pb):
            for u, w in zip(x, y):
                assert np.asarray(u).tobytes() == np.asarray(w).tobytes()
