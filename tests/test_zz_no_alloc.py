"""The allocation-free hull entry point against the allocating one, through
the C ABI: gpu_hull_dandc_dev_noalloc must give gpu_hull_dandc_dev's outputs
bit for bit (vertex and triangle buffers, counts, error words, FP64
volumes), on two calls in a row - the second with a smaller input, so the
shared scratch array still holds the first call's workspaces - and reject
clouds above its size limit with an error."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAX_HV, MAX_HT = 2048, 4096


def _lib():
    lib = C.CDLL(os.path.join(ROOT, "cuacd_amd", "libcuacd_hip.so"))
    lib.gpu_init.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_size_t]
    lib.gpu_destroy.argtypes = [C.c_void_p]
    lib.gpu_last_error.argtypes = [C.c_void_p]
    lib.gpu_last_error.restype = C.c_char_p
    lib.gpu_dev_alloc.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]
    lib.gpu_dev_free.argtypes = [C.c_void_p, C.c_void_p]
    lib.gpu_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    lib.gpu_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    lib.gpu_sync.argtypes = [C.c_void_p]
    for f in (lib.gpu_hull_dandc_dev, lib.gpu_hull_dandc_dev_noalloc):
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6
    return lib


def _ball_clouds(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        out.append((d * rng.uniform(0, 1, (n, 1)) ** (1 / 3)).astype(np.float32))
    return out


class _Dev:
    """Device buffers of one call: inputs uploaded, outputs read back."""

    def __init__(self, lib, h, clouds):
        self.lib, self.h = lib, h
        pts = np.ascontiguousarray(np.concatenate(clouds))
        off = np.zeros(len(clouds) + 1, np.int32)
        off[1:] = np.cumsum([len(c) for c in clouds])
        self.n = len(clouds)
        self.bufs = []
        self.pts = self._up(pts)
        self.off = self._up(off)

    def _alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.gpu_dev_alloc(self.h, max(nbytes, 16), C.byref(p)) == 0
        self.bufs.append(p)
        return p

    def _up(self, a):
        p = self._alloc(a.nbytes)
        assert self.lib.gpu_memcpy_h2d(self.h, p, a.ctypes.data, a.nbytes) == 0
        return p

    def run(self, fn, max_pts):
        n = self.n
        outs = {"verts": np.zeros((n, MAX_HV, 3), np.float32), "tris": np.zeros((n, MAX_HT, 3), np.int32),
                "nv": np.zeros(n, np.int32), "nt": np.zeros(n, np.int32), "err": np.zeros(n, np.int32),
                "vol": np.zeros(n, np.float64)}
        dev = {k: self._alloc(a.nbytes) for k, a in outs.items()}
        rc = fn(self.h, self.pts, self.off, n, max_pts, MAX_HV, MAX_HT, dev["verts"], dev["tris"], dev["nv"],
                dev["nt"], dev["err"], dev["vol"])
        if rc != 0:
            return rc, None
        assert self.lib.gpu_sync(self.h) == 0
        for k, a in outs.items():
            assert self.lib.gpu_memcpy_d2h(self.h, a.ctypes.data, dev[k], a.nbytes) == 0
        assert self.lib.gpu_sync(self.h) == 0
        return 0, outs

    def close(self):
        for p in self.bufs:
            self.lib.gpu_dev_free(self.h, p)
        self.lib.gpu_sync(self.h)


def test_noalloc_hull_matches_the_allocating_entry():
    lib = _lib()
    h = C.c_void_p()
    assert lib.gpu_init(C.byref(h), 0, 64 << 20) == 0, lib.gpu_last_error(None)
    try:
        rng = np.random.default_rng(5)
        # call 1: one-wave clouds and block-tier clouds; at 8000 points 69
        # workspaces fit the scratch, so the block tier runs in several passes
        sizes1 = list(rng.integers(16, 1025, 300)) + list(rng.integers(1025, 8001, 150))
        rng.shuffle(sizes1)
        # call 2: fewer, smaller clouds (the scratch still holds call 1's data)
        sizes2 = list(rng.integers(16, 3001, 60))
        for call, (sizes, seed) in enumerate(((sizes1, 11), (sizes2, 12))):
            print(f"[no_alloc] call {call + 1}: {len(sizes)} clouds, up to {max(sizes)} points", flush=True)
            d = _Dev(lib, h, _ball_clouds(sizes, seed))
            try:
                mp = int(max(sizes))
                rc, want = d.run(lib.gpu_hull_dandc_dev, mp)
                assert rc == 0, lib.gpu_last_error(h)
                rc, got = d.run(lib.gpu_hull_dandc_dev_noalloc, mp)
                assert rc == 0, lib.gpu_last_error(h)
                assert (want["err"] == 0).all() and (want["nv"] >= 4).all()
                for k in ("nv", "nt", "err"):
                    assert np.array_equal(got[k], want[k]), k
                assert got["vol"].tobytes() == want["vol"].tobytes()
                for i in range(len(sizes)):
                    nv, nt = want["nv"][i], want["nt"][i]
                    assert got["verts"][i, :nv].tobytes() == want["verts"][i, :nv].tobytes(), i
                    assert np.array_equal(got["tris"][i, :nt], want["tris"][i, :nt]), i
            finally:
                d.close()
        # above the size limit: an error, nothing launched
        d = _Dev(lib, h, _ball_clouds([20000], 13))
        try:
            rc, _ = d.run(lib.gpu_hull_dandc_dev_noalloc, 20000)
            assert rc != 0 and b"16384" in lib.gpu_last_error(h)
        finally:
            d.close()
    finally:
        lib.gpu_destroy(h)
