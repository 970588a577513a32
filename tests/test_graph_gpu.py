"""GPU test of the Hausdorff plan's HIP graph, called through the C-ABI.

gpu_hausdorff_plan_create captures the preparation and query launches as a
graph; every gpu_hausdorff_plan_run replays it. Each replay (with eager work
and synchronisations of the program in between) must reproduce the eager
result of gpu_hausdorff_batch bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

from cuacd_amd import datasets

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pack(ms):
    vs = [np.asarray(v, np.float32).reshape(-1, 3) for v, _ in ms]
    ts = [np.asarray(t, np.int32).reshape(-1, 3) for _, t in ms]
    voff = np.zeros(len(ms) + 1, np.int32)
    toff = np.zeros(len(ms) + 1, np.int32)
    voff[1:] = np.cumsum([len(v) for v in vs])
    toff[1:] = np.cumsum([len(t) for t in ts])
    return np.ascontiguousarray(np.concatenate(vs)), voff, np.ascontiguousarray(np.concatenate(ts)), toff


def test_plan_graph_replays_match_eager(ctx):
    lib = C.CDLL(os.path.join(ROOT, "cuacd_amd", "libcuacd_hip.so"))
    lib.gpu_init.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_size_t]
    lib.gpu_dev_alloc.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]
    lib.gpu_dev_free.argtypes = [C.c_void_p, C.c_void_p]
    lib.gpu_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    lib.gpu_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    lib.gpu_memset_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]
    lib.gpu_sync.argtypes = [C.c_void_p]
    lib.gpu_hausdorff_plan_create.argtypes = [C.c_void_p] + [C.c_void_p] * 8 + [C.c_int, C.c_int, C.c_int] + \
        [C.c_void_p] * 4 + [C.POINTER(C.c_void_p)]
    lib.gpu_hausdorff_plan_run.argtypes = [C.c_void_p, C.c_void_p]
    lib.gpu_hausdorff_plan_destroy.argtypes = [C.c_void_p, C.c_void_p]
    lib.gpu_destroy.argtypes = [C.c_void_p]
    lib.gpu_last_error.argtypes = [C.c_void_p]
    lib.gpu_last_error.restype = C.c_char_p

    a = datasets.noisy_icospheres(24, 3, 0.15, seed=5)
    b = datasets.noisy_icospheres(24, 2, 0.05, seed=6)
    n = len(a)
    eager_ab, eager_ba = ctx.batch_hausdorff(a, b, sampler="reference")

    g = C.c_void_p()
    assert lib.gpu_init(C.byref(g), 0, C.c_size_t(1 << 28)) == 0
    bufs = []
    try:
        def up(arr):
            p = C.c_void_p()
            assert lib.gpu_dev_alloc(g, C.c_size_t(max(arr.nbytes, 4)), C.byref(p)) == 0
            bufs.append(p)
            assert lib.gpu_memcpy_h2d(g, p, arr.ctypes.data, C.c_size_t(arr.nbytes)) == 0
            return p

        av, avo, at, ato = _pack(a)
        bv, bvo, bt, bto = _pack(b)
        d_av, d_at, d_bv, d_bt = up(av), up(at), up(bv), up(bt)
        out = np.full(2 * n, 7.0, np.float32)
        d_out = up(out)  # stale values: the kernels clear their own outputs
        d_ab = C.c_void_p(d_out.value)
        d_ba = C.c_void_p(d_out.value + 4 * n)
        plan = C.c_void_p()
        rc = lib.gpu_hausdorff_plan_create(g, d_av, d_at, d_bv, d_bt, avo.ctypes.data, ato.ctypes.data,
                                           bvo.ctypes.data, bto.ctypes.data, n, 0, 0, d_ab, d_ba, None, None,
                                           C.byref(plan))
        assert rc == 0, lib.gpu_last_error(g)
        try:
            for rep in range(4):
                assert lib.gpu_hausdorff_plan_run(g, plan) == 0, lib.gpu_last_error(g)
                res = np.empty(2 * n, np.float32)
                assert lib.gpu_memcpy_d2h(g, res.ctypes.data, d_out, C.c_size_t(res.nbytes)) == 0
                assert np.array_equal(np.sqrt(res[:n]), eager_ab), rep
                assert np.array_equal(np.sqrt(res[n:]), eager_ba), rep
                # eager work of the program between replays
                e_ab, e_ba = ctx.batch_hausdorff(a[:3], b[:3], sampler="reference")
                assert np.array_equal(e_ab, eager_ab[:3])
                assert lib.gpu_memset_dev(g, d_out, 0x7f, C.c_size_t(res.nbytes)) == 0
                assert lib.gpu_sync(g) == 0
        finally:
            lib.gpu_hausdorff_plan_destroy(g, plan)
    finally:
        for p in bufs:
            lib.gpu_dev_free(g, p)
        lib.gpu_destroy(g)


def test_score_plan_replays_match_eager_scoring(ctx):
    """The scorer as one device-sized level, called through the C-ABI:
    gpu_score_plan_create sizes every stage on the device (cut, hull tasks,
    k-DOP prefilter, tier lists, Hausdorff tasks and workspace, frees) and
    captures the launches as a HIP graph; every gpu_score_plan_run replays it.
    Each replay - with eager scoring of another batch in between - gives the
    scores and detail of gpu_score_cuts bit for bit, and the device heap is
    back where it started once the plan is gone."""
    from cuacd_amd import sdf
    lib = C.CDLL(os.path.join(ROOT, "cuacd_amd", "libcuacd_hip.so"))
    lib.gpu_init.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_size_t]
    lib.gpu_destroy.argtypes = [C.c_void_p]
    lib.gpu_last_error.argtypes = [C.c_void_p]
    lib.gpu_last_error.restype = C.c_char_p
    lib.gpu_heap_stats.argtypes = [C.c_void_p, C.c_void_p]
    lib.gpu_score_cuts.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                                                      C.c_float] + [C.c_void_p] * 3
    lib.gpu_score_plan_create.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int, C.c_void_p, C.c_void_p,
                                                                             C.c_int, C.c_float,
                                                                             C.POINTER(C.c_void_p)]
    lib.gpu_score_plan_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.gpu_score_plan_destroy.argtypes = [C.c_void_p, C.c_void_p]
    lib.gpu_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.gpu_get_tuning.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]

    meshes = sdf.sdf_meshes(6, seed=21, res=32) + datasets.noisy_icospheres(4, 3, 0.15, seed=3)
    rng = np.random.default_rng(9)
    planes, mesh_of = [], []
    for m, (v, t) in enumerate(meshes):
        for _ in range(8):
            nrm = rng.standard_normal(3)
            nrm /= np.linalg.norm(nrm)
            c = v.mean(0) + rng.uniform(-0.3, 0.3, 3) * (v.max(0) - v.min(0))
            planes.append([*nrm, -float(nrm @ c)])
            mesh_of.append(m)
    planes = np.array(planes, np.float32)
    mesh_of = np.array(mesh_of, np.int32)
    n = len(mesh_of)
    allv, voff, allt, toff = _pack(meshes)
    tau = 0.05

    def eager(g, sel):
        mo = np.ascontiguousarray(mesh_of[sel])
        pl = np.ascontiguousarray(planes[sel])
        sc = np.zeros(len(mo), np.float32)
        dt = np.zeros((len(mo), 10), np.float64)
        rc = lib.gpu_score_cuts(g, allv.ctypes.data, voff.ctypes.data, allt.ctypes.data, toff.ctypes.data,
                                len(meshes), mo.ctypes.data, pl.ctypes.data, len(mo), tau, sc.ctypes.data,
                                dt.ctypes.data, None)
        assert rc == 0, lib.gpu_last_error(g)
        return sc, dt

    g = C.c_void_p()
    assert lib.gpu_init(C.byref(g), 0, C.c_size_t(1 << 30)) == 0
    chunk_above = C.c_int()
    assert lib.gpu_get_tuning(g, 4, C.byref(chunk_above)) == 0
    try:
        _check_plan(lib, g, eager, allv, voff, allt, toff, len(meshes), mesh_of, planes, n, tau)
        # the oversized-set path: with chunk_above at 1024 the halves above it
        # are hulled in chunks (on the device in the plan, by the host's loop
        # in the eager run) and must still give the same scores
        assert lib.gpu_set_tuning(g, 4, 1024) == 0
        _check_plan(lib, g, eager, allv, voff, allt, toff, len(meshes), mesh_of, planes, n, tau)
    finally:
        lib.gpu_set_tuning(g, 4, chunk_above.value)
        lib.gpu_destroy(g)


def _check_plan(lib, g, eager, allv, voff, allt, toff, n_meshes, mesh_of, planes, n, tau):
    """A plan of the whole batch: three replays equal to the eager scores and
    detail, eager work between them, and the heap balanced afterwards."""
    h0 = np.zeros(6, np.uint64)
    assert lib.gpu_heap_stats(g, h0.ctypes.data) == 0
    want_s, want_d = eager(g, slice(None))
    assert np.isfinite(want_s).sum() > n // 2 and (want_d[:, 7:9] > 0).any()  # some halves measure H
    plan = C.c_void_p()
    rc = lib.gpu_score_plan_create(g, allv.ctypes.data, voff.ctypes.data, allt.ctypes.data, toff.ctypes.data,
                                   n_meshes, mesh_of.ctypes.data, planes.ctypes.data, n, tau, C.byref(plan))
    assert rc == 0, lib.gpu_last_error(g)
    try:
        for rep in range(3):
            sc = np.full(n, -7.0, np.float32)
            dt = np.full((n, 10), -7.0, np.float64)
            rc = lib.gpu_score_plan_run(g, plan, sc.ctypes.data, dt.ctypes.data)
            assert rc == 0, lib.gpu_last_error(g)
            assert sc.tobytes() == want_s.tobytes(), rep
            assert dt.tobytes() == want_d.tobytes(), rep
            # eager scoring of another batch between the replays
            s2, _ = eager(g, slice(0, n, 3))
            assert s2.tobytes() == want_s[0:n:3].tobytes()
    finally:
        lib.gpu_score_plan_destroy(g, plan)
    h1 = np.zeros(6, np.uint64)
    assert lib.gpu_heap_stats(g, h1.ctypes.data) == 0
    assert h1[0] == h0[0] and h1[3] == h0[3], (h0, h1)
