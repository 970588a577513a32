"""GPU tests of the batched exact cut scorer (Context.batch_score_cuts).

The scorer is the lookahead's level-0 work per (part, plane) pair - full cut
(reference cuda/plane_cut.cuh), exact hulls of both halves, the stop cost
max(0.3 rv, H) of each half with H measured when 0.3 rv < tau
(cuda/la_common.cuh:21-34, cuda/la_refine.cu:19-45). The serial oracle
(lookahead_oracle.cut_cost) composes the C oracle's cut, exact hull and the
reference-sampler Hausdorff distance with exact FP64 distances.
"""
import numpy as np
import pytest

from cuacd_amd import datasets, sdf

pytestmark = pytest.mark.gpu


def _cases():
    meshes = sdf.sdf_meshes(3, seed=31, res=18)
    out = []
    for m, (v, t) in enumerate(meshes):
        pl = datasets.cut_candidates(v, t, n_edges=6, n_total=12, seed=m)
        out.append((v, t, pl))
    return out


def _check(ctx, tau):
    import lookahead_oracle as LO
    cases = _cases()
    meshes = [(v, t) for v, t, _ in cases]
    mesh_of = np.concatenate([np.full(len(p), m, np.int32) for m, (_, _, p) in enumerate(cases)])
    planes = np.concatenate([p for _, _, p in cases])
    score, detail, stats = ctx.batch_score_cuts(meshes, mesh_of, planes, threshold=tau, return_detail=True)
    assert stats[5] > 0
    n_h = 0
    for i in range(len(planes)):
        v, t = meshes[mesh_of[i]]
        ref, halves = LO.cut_cost(v, t, planes[i], tau)
        if halves is None:
            assert score[i] == np.float32(1e30) and detail[i, 9] == 0, i
            continue
        assert detail[i, 9] == 1, i
        for s, (mv, hv, crv, h) in enumerate(halves):
            assert abs(detail[i, 1 + s] - mv) <= 1e-9 * max(abs(mv), 1e-12), (i, s)
            assert abs(detail[i, 3 + s] - hv) <= 1e-6 * max(hv, 1e-12), (i, s)
            assert abs(detail[i, 5 + s] - crv) <= 1e-5 * max(crv, 1e-6), (i, s)
            if h >= 0:
                n_h += 1
                # Hausdorff to 1e-5 absolute (the reference's test tolerance)
                assert abs(detail[i, 7 + s] - h) <= 1e-5, (i, s, detail[i, 7 + s], h)
            else:
                assert detail[i, 7 + s] == -1.0, (i, s)
        assert abs(score[i] - ref) <= 1e-5 + 1e-5 * ref, (i, score[i], ref)
    return n_h


def test_scores_match_oracle(ctx):
    _check(ctx, 0.05)


def test_scores_match_oracle_all_hausdorff(ctx):
    # a threshold above every volume cost measures H on every half
    assert _check(ctx, 10.0) > 0


def test_scores_match_oracle_chunked_hulls(ctx):
    # sets above 1024 points after the prefilter are hulled in chunks of 512
    # first; the union of the chunk corners must give the same hulls
    import lookahead_oracle as LO
    old = ctx.get_tuning("hull_chunk_above")
    try:
        ctx.set_tuning("hull_chunk_above", 1024)
        assert ctx.get_tuning("hull_chunk_above") == 1024
        _check(ctx, 10.0)
        # a nearly spherical mesh: halves of ~5k points that nearly all pass
        # the prefilter, several chunking rounds
        (v, t), = datasets.noisy_icospheres(1, 5, 0.003, seed=2)
        v = np.asarray(v, np.float32)
        planes = np.array([[1, 0, 0, -0.1], [0.6, 0.8, 0, 0.2]], np.float32)
        score, detail, _ = ctx.batch_score_cuts([(v, t)], [0, 0], planes, threshold=0.05, return_detail=True)
        for i, pl in enumerate(planes):
            ref, halves = LO.cut_cost(v, t, pl, 0.05)
            for s_, (mv, hv, crv, h) in enumerate(halves):
                assert abs(detail[i, 3 + s_] - hv) <= 1e-6 * hv, (i, s_)
                assert abs(detail[i, 7 + s_] - h) <= 1e-5, (i, s_)
            assert abs(score[i] - ref) <= 1e-5, (i, score[i], ref)
    finally:
        ctx.set_tuning("hull_chunk_above", old)


def test_missed_plane_and_empty_batch(ctx):
    v, t = datasets.cube_mesh()
    v = np.asarray(v, np.float32)
    planes = np.array([[1, 0, 0, -5.0], [1, 0, 0, -0.5]], np.float32)
    score, detail, _ = ctx.batch_score_cuts([(v, t)], [0, 0], planes, threshold=0.05, return_detail=True)
    assert score[0] == np.float32(1e30) and detail[0, 0] == 2
    # two boxes: both halves convex, cost 0
    assert detail[1, 9] == 1 and score[1] <= 1e-6
    assert len(ctx.batch_score_cuts([(v, t)], [], np.zeros((0, 4), np.float32))) == 0


def test_heap_balanced_after_scoring(ctx):
    cases = _cases()[:1]
    v, t, pl = cases[0]
    before = ctx.heap_stats()
    ctx.batch_score_cuts([(v, t)], np.zeros(len(pl), np.int32), pl, threshold=0.05)
    after = ctx.heap_stats()
    assert after[0] == before[0] and after[3] == before[3], (before, after)


def test_autotuner_keeps_scores(ctx):
    from cuacd_amd import autotune
    cases = _cases()
    meshes = [(v, t) for v, t, _ in cases]
    mesh_of = np.concatenate([np.full(len(p), m, np.int32) for m, (_, _, p) in enumerate(cases)])
    planes = np.concatenate([p for _, _, p in cases])
    start = ctx.get_tuning()
    ref = ctx.batch_score_cuts(meshes, mesh_of, planes, threshold=10.0)
    try:
        best, trials = autotune.tune(ctx, lambda: ctx.batch_score_cuts(meshes, mesh_of, planes, threshold=10.0))
        # every knob only picks kernels: no setting may change a score
        assert all(ok for _, _, ok in trials), [s for s, _, ok in trials if not ok]
        assert np.array_equal(ctx.batch_score_cuts(meshes, mesh_of, planes, threshold=10.0), ref)
    finally:
        for k, v in start.items():
            ctx.set_tuning(k, v)


def test_launch_knobs_keep_scores_on_large_halves(ctx):
    """Halves of 64^3 SDF meshes reach every hull tier (257-16384 points after
    the prefilter): the launch order, the eight-wave wave tiers, multi-apex
    rounds in the LDS and the HBM-workspace tiers pick kernels only - every
    score and every half's hull volume must be the same bits."""
    meshes = sdf.sdf_meshes(4, seed=5, res=64, min_tris=12000, max_tris=60000)
    planes = [datasets.cut_candidates(v, t, n_edges=6, n_total=12, seed=m) for m, (v, t) in enumerate(meshes)]
    mesh_of = np.concatenate([np.full(len(p), m, np.int32) for m, p in enumerate(planes)])
    planes = np.concatenate(planes)
    start = ctx.get_tuning()
    import cuacd_amd
    ref_s, ref_d, st = ctx.batch_score_cuts(meshes, mesh_of, planes, threshold=0.05, return_detail=True)
    launches = {k: st[6 + 3 * i + 1] for i, k in enumerate(cuacd_amd.LOOKAHEAD_KERNELS)}
    # the multi-wave tiers the knobs act on were used
    assert launches["hull_w1024"] > 0 and launches["hull_b2048"] > 0 and launches["hull_b4096"] > 0, launches
    variants = [{"hull_order": 0}, {"hull_w8": 0}, {"hull_w8": 15}, {"hull_multi_hbm": 0}, {"hull_multi": 0},
                {"hull_multi": 0, "hull_multi_hbm": 0, "hull_w8": 0, "hull_order": 0}]
    try:
        for var in variants:
            for k, v in var.items():
                ctx.set_tuning(k, v)
            s, d, _ = ctx.batch_score_cuts(meshes, mesh_of, planes, threshold=0.05, return_detail=True)
            assert np.array_equal(s, ref_s), var
            assert np.array_equal(d, ref_d), var
            for k, v in start.items():
                ctx.set_tuning(k, v)
    finally:
        for k, v in start.items():
            ctx.set_tuning(k, v)
