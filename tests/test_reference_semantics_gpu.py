"""The reference test suite's remaining semantics, on the GPU:

* the record sort of test_warp_sort (reference tests/test_warp_sort.py:13-17,
  69-142), including arrays longer than one workgroup's LDS run;
* GPU mesh volumes against the FP64 CPU volume (reference
  tests/test_hull_mesh.py:107-123);
* the 49160.stl decomposition (reference tests/test_lookahead.py:177-187),
  here also checked part for part against the serial oracle;
* the width check (reference tests/test_concave_edges.py:100-106);
* a device pool too small for the search: an error, not a crash, and no
  block left allocated (reference SURVEY of the allocator's OOM path).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from cuacd_amd import datasets, meshio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STL_49160 = os.path.join(ROOT, "tests", "golden", "49160.stl")


def _lib():
    lib = C.CDLL(os.path.join(ROOT, "cuacd_amd", "libcuacd_hip.so"))
    lib.gpu_init.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_size_t]
    lib.gpu_destroy.argtypes = [C.c_void_p]
    lib.gpu_last_error.argtypes = [C.c_void_p]
    lib.gpu_last_error.restype = C.c_char_p
    lib.gpu_heap_stats.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong)]
    lib.gpu_test_warp_sort.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    lib.gpu_test_mesh_volume.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    lib.gpu_batch_mesh_volume.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_int, C.c_void_p]
    return lib


@pytest.fixture(scope="module")
def abi(ctx):
    """A second context through the C ABI (the Python module keeps its own)."""
    lib = _lib()
    h = C.c_void_p()
    assert lib.gpu_init(C.byref(h), 0, 256 << 20) == 0, lib.gpu_last_error(None)
    yield lib, h
    lib.gpu_destroy(h)


def _p(a):
    return a.ctypes.data


# ---- record sort -------------------------------------------------------------

def _key_sorted(arr):
    """(y, x, z, index) order (reference ws_less)."""
    return arr[np.lexsort((arr[:, 3], arr[:, 2], arr[:, 0], arr[:, 1]))]


def _random_records(n, rng):
    a = np.empty((n, 4), np.int32)
    a[:, :3] = rng.integers(-5108, 5109, size=(n, 3))
    a[:, 3] = np.arange(n)
    return a


def _gpu_sort(abi, arrays):
    lib, h = abi
    off = np.zeros(len(arrays) + 1, np.int32)
    off[1:] = np.cumsum([len(a) for a in arrays])
    flat = np.ascontiguousarray(np.concatenate(arrays).astype(np.int32))
    assert lib.gpu_test_warp_sort(h, _p(flat), int(off[-1]), _p(off), len(arrays)) == 0, lib.gpu_last_error(h)
    return [flat[off[i]:off[i + 1]] for i in range(len(arrays))]


@pytest.mark.parametrize("n", [1, 2, 5, 16, 31, 32, 33, 50, 100, 200, 500, 1000, 2000, 4096, 4097, 12289, 70000])
def test_sort_sizes(abi, n):
    arr = _random_records(n, np.random.default_rng(42 + n))
    [got] = _gpu_sort(abi, [arr])
    np.testing.assert_array_equal(got, _key_sorted(arr))


def test_sort_batch_and_special_orders(abi):
    rng = np.random.default_rng(99)
    arrays = [_random_records(s, rng) for s in (10, 32, 33, 100, 500, 9000, 0, 3)]
    asc = np.zeros((64, 4), np.int32)
    asc[:, 1] = np.arange(64)
    asc[:, 3] = np.arange(64)
    desc = np.zeros((128, 4), np.int32)
    desc[:, 1] = np.arange(128)[::-1]
    desc[:, 3] = np.arange(128)
    equal = np.zeros((100, 4), np.int32)
    equal[:, 3] = np.arange(100)
    dup = np.empty((200, 4), np.int32)
    dup[:, :3] = rng.integers(0, 3, size=(200, 3))
    dup[:, 3] = np.arange(200)
    arrays += [asc, desc, equal, dup]
    for a, got in zip(arrays, _gpu_sort(abi, arrays)):
        np.testing.assert_array_equal(got, _key_sorted(a))


# ---- mesh volumes ------------------------------------------------------------

def _cube_points():
    return np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32)


def test_mesh_volume_of_the_cube_hull(ctx, abi):
    lib, h = abi
    (hv, ht, vol), = ctx.batch_hull_dandc_mesh([_cube_points()])
    assert len(hv) == 8 and len(ht) == 12
    out = np.zeros(1, np.float32)
    hv = np.ascontiguousarray(hv, np.float32)
    ht = np.ascontiguousarray(ht, np.int32)
    assert lib.gpu_test_mesh_volume(h, _p(hv), len(hv), _p(ht), len(ht), _p(out)) == 0
    cpu = abs(oracle.mesh_volume(hv, ht))
    assert abs(cpu - 1.0) < 1e-12
    assert abs(out[0] - cpu) <= 1e-7 * cpu


def test_batch_mesh_volume_matches_fp64(ctx, abi):
    lib, h = abi
    meshes = datasets.noisy_icospheres(6, 3, 0.15, seed=4) + [datasets.l_shape_union(), datasets.hollow_tube()]
    vs = [np.asarray(v, np.float32) for v, _ in meshes]
    ts = [np.asarray(t, np.int32) for _, t in meshes]
    voff = np.zeros(len(meshes) + 1, np.int32)
    toff = np.zeros(len(meshes) + 1, np.int32)
    voff[1:] = np.cumsum([len(v) for v in vs])
    toff[1:] = np.cumsum([len(t) for t in ts])
    V = np.ascontiguousarray(np.concatenate(vs))
    T = np.ascontiguousarray(np.concatenate(ts))
    out = np.zeros(len(meshes), np.float32)
    assert lib.gpu_batch_mesh_volume(h, _p(V), len(V), _p(T), len(T), _p(toff), _p(voff), len(meshes), _p(out)) == 0
    for m, (v, t) in enumerate(zip(vs, ts)):
        cpu = abs(oracle.mesh_volume(v, t))
        # FP64 accumulation rounded once to float
        assert abs(out[m] - cpu) <= 1e-7 * cpu, (m, out[m], cpu)
    # the Context method gives the same numbers
    assert np.array_equal(np.asarray(ctx.batch_mesh_volume(vs, ts), np.float32), out)


# ---- 49160.stl -----------------------------------------------------------------

def _normalized_49160():
    v, t = meshio.read_stl(STL_49160)
    lo, hi = v.min(0), v.max(0)
    return ((v - (lo + hi) / 2) / ((hi - lo).max() / 2)).astype(np.float32), t


def test_49160_decomposes(ctx):
    """Reference defaults (width 60, 32 concave edges), tau 0.05: at least one
    part, every part closed, the volume kept, every part under tau."""
    v, t = _normalized_49160()
    assert oracle.edge_manifold_report(t) == (0, 0)
    parts, info = ctx.batch_lookahead_decompose([(v, t)], return_info=True, max_iters=100, threshold=0.05)
    parts, info = parts[0], info[0]
    assert len(parts) >= 1
    total = abs(oracle.mesh_volume(v, t))
    assert abs(sum(abs(oracle.mesh_volume(p[0], p[1])) for p in parts) - total) <= 1e-4 * total
    for p in parts:
        assert oracle.edge_manifold_report(p[1]) == (0, 0)
    rv = 0.3 * np.cbrt(3 / (4 * np.pi) * np.maximum(info["hull_vol"] - info["mesh_vol"], 0))
    assert info["iters"] < 100 and (np.maximum(rv, info["hausdorff"]) < 0.05).all()


def test_49160_parity_with_serial_oracle_merge_on(ctx):
    """The command-line defaults (width/width2 30/5, 16 concave edges, 16
    parts cut per iteration, hull merge on) to convergence: the GPU search and
    the serial oracle give the same parts, bit for bit."""
    import lookahead_oracle as LO
    v, t = _normalized_49160()
    kw = dict(max_iters=100, width=30, width2=5, threshold=0.05, depth=2, n_concave_edges=16, max_n_cutting=16,
              merge_hulls=True)
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
        assert ghv.tobytes() == rp[2].tobytes() and np.array_equal(ght, rp[3])


# ---- argument checks and the out-of-memory path ----------------------------------

def test_width_validation(ctx):
    v, t = datasets.l_shape_union()
    with pytest.raises(ValueError, match="must be < 512"):
        ctx.lookahead_decompose(v, t, width=500, n_concave_edges=5)


class _Part(C.Structure):
    _fields_ = [("verts", C.c_void_p), ("tris", C.c_void_p), ("nv", C.c_int), ("nt", C.c_int),
                ("hull_verts", C.c_void_p), ("hull_tris", C.c_void_p), ("hull_nv", C.c_int), ("hull_nt", C.c_int),
                ("mesh_vol", C.c_float), ("hull_vol", C.c_float), ("hausdorff", C.c_float)]


class _Result(C.Structure):
    _fields_ = [("parts", C.POINTER(_Part)), ("nparts", C.c_int)]


def test_tiny_pool_reports_out_of_memory(ctx):
    """A 2 MiB device pool holds a small search (the L-shape decomposes) but
    not a 20k-triangle search's halves: that call fails with the heap's
    out-of-memory error and frees every block it took."""
    lib = _lib()
    lib.lookahead_decompose.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                        C.c_int, C.c_int, C.POINTER(_Result)]
    lib.gpu_result_free.argtypes = [C.POINTER(_Result)]
    from cuacd_amd import sdf
    h = C.c_void_p()
    assert lib.gpu_init(C.byref(h), 0, 2 << 20) == 0, lib.gpu_last_error(None)
    try:
        def run(v, t):
            v = np.ascontiguousarray(v, np.float32)
            t = np.ascontiguousarray(t, np.int32)
            hv, ht = ctx._input_hull(v)
            hv = np.ascontiguousarray(hv, np.float32)
            ht = np.ascontiguousarray(ht, np.int32)
            res = _Result()
            rc = lib.lookahead_decompose(h, _p(v), len(v), _p(t), len(t), _p(hv), len(hv), _p(ht), len(ht),
                                         20, 30, 5, 0.05, 2, 0, 16, 0, 0, 0, 1, 16, 0.005, 3.49, 10, 0,
                                         C.byref(res))
            return rc, res

        v, t = datasets.l_shape_union()
        rc, res = run(v, t)
        assert rc == 0, lib.gpu_last_error(h)
        assert res.nparts >= 2
        lib.gpu_result_free(C.byref(res))
        v, t = sdf.sdf_meshes(1, seed=4, res=48, min_tris=15000, max_tris=40000)[0]
        rc, res = run(v, t)
        assert rc != 0
        assert b"heap:oom" in lib.gpu_last_error(h), lib.gpu_last_error(h)
        st = (C.c_ulonglong * 6)()
        assert lib.gpu_heap_stats(h, st) == 0
        assert st[0] == 0 and st[3] == 0, list(st)  # nothing outstanding on either heap
    finally:
        lib.gpu_destroy(h)
