"""GPU tests of the learned cut proposal (cuacd_amd/csrc/mlp.hip): the BF16
MFMA GEMM against numpy, the network's forward pass and gradients against the
FP64 oracle with the same BF16 rounding points (oracle/mlp_oracle.py), that
training lowers the loss, and the learned search itself."""
import numpy as np
import pytest

import cuacd_amd
from cuacd_amd import proposal as PR

pytestmark = pytest.mark.gpu


def _bf16_bits(x):
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def _bits_to_f(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("M,N,K,ta,tb", [(16, 16, 16, 0, 0), (64, 64, 32, 0, 0), (100, 70, 45, 0, 0),
                                         (130, 128, 144, 1, 0), (128, 16, 300, 0, 1), (33, 129, 17, 1, 1),
                                         (128, 128, 9000, 1, 0), (16, 128, 5000, 1, 0)])
def test_gemm_matches_numpy(ctx, M, N, K, ta, tb):
    rng = np.random.default_rng(M * 7 + N)
    A = _bf16_bits(rng.normal(size=(K, M) if ta else (M, K)))
    B = _bf16_bits(rng.normal(size=(N, K) if tb else (K, N)))
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    C = np.empty((M, N), np.float32)
    cuacd_amd._gpu.mlp_gemm_test(M, N, K, A.ctypes.data, sam, sak, A.size, B.ctypes.data, sbk, sbn, B.size,
                                 C.ctypes.data)
    a = _bits_to_f(A)
    b = _bits_to_f(B)
    ref = (a.T if ta else a) @ (b.T if tb else b)
    assert np.allclose(C, ref, rtol=1e-5, atol=1e-4 * np.sqrt(K)), np.abs(C - ref).max()


def _batch(seed=0, P=3, C=20):
    import mlp_oracle as MO
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-1, 1, (P, MO.NS, 3)).astype(np.float32)
    planes = rng.normal(size=(C, 4)).astype(np.float32)
    planes[:, :3] /= np.linalg.norm(planes[:, :3], axis=1, keepdims=True)
    part_of = np.sort(rng.integers(0, P, C)).astype(np.int32)
    t = rng.uniform(0, 0.5, C).astype(np.float32)
    return pts, planes, part_of, t


def test_forward_and_gradients_match_oracle(ctx):
    import mlp_oracle as MO
    net = PR.ProposalNet(ctx, seed=5)
    try:
        params = net.params()
        assert len(params) == MO.N_PARAMS
        pts, planes, part_of, t = _batch()
        y = net.predict(pts, planes, part_of)
        loss_ref, g_ref, y_ref = MO.loss_and_grads(params.astype(np.float64), pts, planes, part_of, t)
        assert np.allclose(y, y_ref, rtol=2e-3, atol=2e-3), np.abs(y - y_ref).max()
        loss, g = net.train_step(pts, planes, part_of, t, apply=False, return_grads=True)
        assert abs(loss - loss_ref) <= 1e-3 * loss_ref + 1e-6
        gd = MO.unpack(g.astype(np.float64))
        rd = MO.unpack(g_ref)
        for name, _ in MO.SHAPES:
            a, b = gd[name].reshape(-1), rd[name].reshape(-1)
            nb = np.linalg.norm(b)
            if nb == 0:
                assert np.linalg.norm(a) == 0, name
                continue
            rel = np.linalg.norm(a - b) / nb
            assert rel < 2e-2, (name, rel)
        # apply=False left the weights alone
        assert np.array_equal(net.params(), params)
    finally:
        net.close()


def test_fused_inference_equals_layer_by_layer(ctx):
    """The fused point-layer kernel (tuning knob mlp_fused) gives the scores
    of the layer-by-layer GEMMs bit for bit: same MFMA tiles, k order, bias,
    ReLU, BF16 rounding and pooling. Trained weights, a batch with every part
    cut by many planes (more candidates than workgroups, so each workgroup
    strides over several), planes that leave one side empty, and one
    candidate alone."""
    import mlp_oracle as MO
    net = PR.ProposalNet.load_default(ctx)
    old = ctx.get_tuning()["mlp_fused"]
    try:
        rng = np.random.default_rng(9)
        P, C = 7, 2000
        pts = rng.uniform(-1, 1, size=(P, 256, 3)).astype(np.float32)
        n = rng.normal(size=(C, 3))
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        d = rng.uniform(-0.8, 0.8, size=C)
        d[:50] = 2.0      # every sample on the positive side
        d[50:100] = -2.0  # every sample on the negative side
        planes = np.concatenate([n, d[:, None]], axis=1).astype(np.float32)
        part_of = np.sort(rng.integers(0, P, C)).astype(np.int32)
        out = {}
        for fused in (0, 1):
            ctx.set_tuning("mlp_fused", fused)
            out[fused] = net.predict(pts, planes, part_of) + np.float32(0.0)
        assert np.array_equal(out[0], out[1]), np.abs(out[0] - out[1]).max()
        y_ref, _ = MO.forward(net.params().astype(np.float64), pts, planes[:300], part_of[:300])
        assert np.allclose(out[1][:300], y_ref, rtol=2e-3, atol=2e-3)
        one = net.predict(pts[:1], planes[:1], np.zeros(1, np.int32))
        ctx.set_tuning("mlp_fused", 0)
        assert np.array_equal(one, net.predict(pts[:1], planes[:1], np.zeros(1, np.int32)))
    finally:
        ctx.set_tuning("mlp_fused", old)
        net.close()


def test_training_lowers_the_loss(ctx):
    rng = np.random.default_rng(4)
    P, C = 24, 240
    pts = rng.uniform(-1, 1, (P, 256, 3)).astype(np.float32)
    planes = rng.normal(size=(C, 4)).astype(np.float32)
    planes[:, :3] /= np.linalg.norm(planes[:, :3], axis=1, keepdims=True)
    part_of = np.sort(rng.integers(0, P, C)).astype(np.int32)
    # a learnable target: depends on the plane offset and the part's spread
    spread = pts.std(axis=1).mean(axis=1)
    t = (0.2 * np.abs(planes[:, 3]) + spread[part_of]).astype(np.float32)
    net = PR.ProposalNet(ctx, seed=1)
    try:
        l0 = net.train_step(pts, planes, part_of, t, apply=False)
        for _ in range(150):
            loss = net.train_step(pts, planes, part_of, t, lr=2e-3)
        assert loss < 0.2 * l0, (l0, loss)
    finally:
        net.close()


def test_samples_lie_on_the_part_in_its_frame(ctx):
    from cuacd_amd import datasets
    v, t = datasets.box_mesh((0, 0, 0), (2, 1, 1))
    s = PR.sample_parts(ctx, [(v, t)])[0]
    # frame: center (1, .5, .5), scale 1 -> x in [-1, 1], y, z in [-.5, .5], on the box surface
    assert s.shape == (256, 3)
    assert np.all(np.abs(s[:, 0]) <= 1 + 1e-6) and np.all(np.abs(s[:, 1:]) <= 0.5 + 1e-6)
    on = (np.isclose(np.abs(s[:, 0]), 1, atol=1e-5) | np.isclose(np.abs(s[:, 1]), 0.5, atol=1e-5) |
          np.isclose(np.abs(s[:, 2]), 0.5, atol=1e-5))
    assert on.all()
    # area-stratified: the two 1x1 end faces hold ~2/10 of the area
    ends = np.isclose(np.abs(s[:, 0]), 1, atol=1e-5).mean()
    assert 0.14 < ends < 0.26, ends


SEARCH = dict(max_iters=100, width=30, width2=5, threshold=0.05, depth=2, max_n_cutting=16, n_concave_edges=16)


def _cost(info):
    rv = 0.3 * np.cbrt(3 / (4 * np.pi) * np.maximum(info["hull_vol"] - info["mesh_vol"], 0))
    return np.maximum(rv, info["hausdorff"])


def test_learned_search_keeps_quality_with_fewer_exact_evaluations(ctx):
    """The shipped network in the search: every part still meets tau (the
    exact pipeline judges every cut), volumes add up, and level 0 evaluates
    exactly only the kept candidates."""
    import oracle
    from cuacd_amd import sdf
    net = PR.ProposalNet.load_default(ctx)
    try:
        meshes = sdf.sdf_meshes(4, seed=77, res=24)
        plain, pinfo = ctx.batch_lookahead_decompose(meshes, return_info=True, **SEARCH)
        learned, linfo = ctx.batch_lookahead_decompose(meshes, return_info=True, proposal=net, top_k=16, **SEARCH)
        pr = ctx.last_lookahead_proposal
        assert pr["ranked"] > 0 and 0 < pr["kept"] <= 0.5 * pr["ranked"], pr
        for (v, t), parts, info in zip(meshes, learned, linfo):
            assert info["iters"] < SEARCH["max_iters"]
            assert _cost(info).max() <= SEARCH["threshold"] + 1e-6
            total = abs(oracle.mesh_volume(v, t))
            vol = sum(abs(oracle.mesh_volume(pv, pt)) for pv, pt, _, _ in parts)
            assert abs(vol - total) <= 1e-4 * total
        # the plain search is untouched by the proposal's presence elsewhere
        again = ctx.batch_lookahead_decompose(meshes, **SEARCH)
        assert [len(p) for p in again] == [len(p) for p in plain]
    finally:
        net.close()


def test_trained_network_ranks_better_than_untrained(ctx):
    """What training gains: on parts of held-out meshes, the exact best
    level-0 candidate is in the trained network's top 8 more often, and the
    kept candidates' best exact score is closer to the true best (regret),
    than with the untrained initialisation."""
    from cuacd_amd import sdf
    meshes = sdf.sdf_meshes(6, seed=99, res=32)
    parts = PR.training_parts(ctx, meshes, iters=(0, 1), **{k: v for k, v in SEARCH.items() if k != "max_iters"})
    data = PR.make_dataset(ctx, parts, threshold=0.05, width=30, n_concave_edges=16)
    trained = PR.ProposalNet.load_default(ctx)
    untrained = PR.ProposalNet(ctx, seed=3)
    try:
        hit_t, reg_t, n = PR.ranking_quality(trained, data, top_k=8)
        hit_u, reg_u, _ = PR.ranking_quality(untrained, data, top_k=8)
        assert n >= 20
        assert hit_t > hit_u and reg_t < reg_u, (hit_t, hit_u, reg_t, reg_u)
    finally:
        trained.close()
        untrained.close()


def test_search_record_and_error_paths(ctx):
    from cuacd_amd import sdf
    meshes = sdf.sdf_meshes(2, seed=13, res=20)
    pts, pl, po, tg = PR.record_dataset(ctx, meshes, **dict(SEARCH, max_iters=2, merge_hulls=False))
    assert len(pts) > 0 and len(pl) == len(po) == len(tg) and pts.shape[1:] == (256, 3)
    assert np.all(np.diff(po) >= 0) and po.max() == len(pts) - 1
    assert np.all(np.abs(np.linalg.norm(pl[:, :3], axis=1) - 1) < 1e-5)
    # path costs are costs (>= 0) or the failure score; every part has a finite best
    assert np.all(tg >= 0) and np.all(tg <= PR.FAIL_SCORE)
    starts = np.searchsorted(po, np.arange(len(pts) + 1))
    assert all(tg[a:b].min() < PR.FAIL_SCORE for a, b in zip(starts[:-1], starts[1:]))
    # not recording: the size query fails loudly
    with pytest.raises(RuntimeError):
        cuacd_amd._gpu.lookahead_record_size()
    # a proposal needs a positive top_k
    net = PR.ProposalNet(ctx, seed=0)
    try:
        with pytest.raises(RuntimeError):
            ctx.batch_lookahead_decompose(meshes[:1], proposal=net, top_k=0, **dict(SEARCH, max_iters=1))
        with pytest.raises(ValueError):
            net.predict(np.zeros((2, 256, 3), np.float32), np.zeros((2, 4), np.float32), [1, 0])
    finally:
        net.close()
    # the failed call left no proposal behind: the plain search still runs
    ctx.batch_lookahead_decompose(meshes[:1], **dict(SEARCH, max_iters=1))


def test_recorded_targets_are_every_first_cuts_best_path(ctx):
    """The training record holds, per cutting part, every level-0
    candidate's best path cost - the oracle's per-first-cut minimum over all
    leaves. The search's two-round leaf bound may skip leaves that cannot
    change the argmin; a recording run scores them all, so no candidate is
    recorded too high or as failed."""
    import lookahead_oracle as LO
    from cuacd_amd import sdf
    kw = dict(SEARCH, max_iters=3, merge_hulls=False)
    for v, t in sdf.sdf_meshes(2, seed=5, res=20):
        pts, pl, po, tg = PR.record_dataset(ctx, [(v, t)], **kw)
        hv, ht = ctx._input_hull(v)
        trace = []
        LO.lookahead(v, t, hv, ht, trace=trace, **kw)
        starts = np.searchsorted(po, np.arange(len(pts) + 1))
        recorded = [np.sort(tg[a:b][tg[a:b] < PR.FAIL_SCORE]) for a, b in zip(starts[:-1], starts[1:])]
        want = [np.sort(np.array([c for c in tr["per_cut"] if c < 1e29], np.float32)) for tr in trace]
        # compare the parts with at least one finite path (a part with fewer
        # than two candidates is not recorded; one whose cuts all failed has
        # nothing to compare)
        recorded = [r for r in recorded if len(r)]
        want = [w for w in want if len(w)]
        assert len(recorded) == len(want)
        for r, w in zip(recorded, want):
            assert len(r) == len(w) and np.allclose(r, w, rtol=0, atol=1e-7), (r[:5], w[:5])


BENCH_SEARCH = dict(max_iters=100, width=30, width2=5, depth=2, max_n_cutting=16, n_concave_edges=16,
                    merge_hulls=True)


@pytest.mark.parametrize("tau,top_k", [(0.05, 16), (0.03, 24)])
def test_learned_search_on_the_lookahead_benchmark_meshes(ctx, tau, top_k):
    """The learned proposal on the 64-mesh lookahead workload (SDF blobs,
    seed 4, 48^3 grid, 2k-50k triangles) at the benchmark's operating points
    (top 16 candidates per part at tau 0.05, top 24 at 0.03): every part of
    every converged mesh meets tau, no mesh is left at max_iters that the
    plain search finished, the part count stays within 10 % of the plain
    search's, and the network really pruned (kept <= a third of the ranked
    candidates)."""
    from cuacd_amd import sdf
    meshes = sdf.sdf_meshes(64, seed=4, res=48, min_tris=2000, max_tris=50000)
    kw = dict(BENCH_SEARCH, threshold=tau)
    net = PR.ProposalNet.load_default(ctx)
    try:
        plain, pinfo = ctx.batch_lookahead_decompose(meshes, return_info=True, **kw)
        learned, linfo = ctx.batch_lookahead_decompose(meshes, return_info=True, proposal=net, top_k=top_k, **kw)
        pr = ctx.last_lookahead_proposal
    finally:
        net.close()
    assert pr["ranked"] > 0 and 0 < pr["kept"] <= pr["ranked"] / 3, pr
    for pi, li in zip(pinfo, linfo):
        if pi["iters"] < kw["max_iters"]:
            assert li["iters"] < kw["max_iters"]
        if li["iters"] < kw["max_iters"]:
            assert _cost(li).max() <= tau + 1e-6
    n_plain = sum(len(p) for p in plain)
    n_learned = sum(len(p) for p in learned)
    assert n_learned <= 1.10 * n_plain, (n_learned, n_plain)
