"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product path (cuacd_amd).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run python cuacd_amd/build.py oracle")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_quantize.argtypes = [P, C.c_int, P]
        L.orc_hull.argtypes = [P, C.c_int, P, P, C.c_int]
        L.orc_hull.restype = C.c_int
        L.orc_general_position.argtypes = [P, C.c_int, P, C.c_int]
        L.orc_general_position.restype = C.c_int
        L.orc_mesh_volume.argtypes = [P, P, C.c_int]
        L.orc_mesh_volume.restype = C.c_double
        L.orc_hull_bruteforce.argtypes = [P, C.c_int, P]
        L.orc_hull_bruteforce.restype = C.c_int
        L.orc_hull_batch.argtypes = [P, P, C.c_int, P, P, P, P, C.c_int]
        L.orc_hull_batch.restype = C.c_int
        L.orc_hull_batch_serial.argtypes = [P, P, C.c_int, P, P, P]
        L.orc_hull_batch_serial.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def quantize(pts):
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    q = np.empty((len(pts), 3), np.int32)
    lib().orc_quantize(_p(pts), len(pts), _p(q))
    return q


def hull(q):
    """Exact hull of integer points: (vertex flags bool[n], faces int[nf,3])."""
    q = np.ascontiguousarray(q, np.int32).reshape(-1, 3)
    n = len(q)
    flags = np.zeros(max(n, 1), np.uint8)
    cap = 2 * n + 8
    faces = np.zeros((cap, 3), np.int32)
    nf = lib().orc_hull(_p(q), n, _p(flags), _p(faces), cap)
    if nf < 0:
        raise RuntimeError("oracle hull failed")
    return flags[:n].astype(bool), faces[:nf].copy()


def general_position(q, faces):
    q = np.ascontiguousarray(q, np.int32)
    faces = np.ascontiguousarray(faces, np.int32)
    return bool(lib().orc_general_position(_p(q), len(q), _p(faces), len(faces)))


def mesh_volume(pts, tris):
    pts = np.ascontiguousarray(pts, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    return lib().orc_mesh_volume(_p(pts), _p(tris), len(tris))


def hull_bruteforce(q):
    q = np.ascontiguousarray(q, np.int32).reshape(-1, 3)
    flags = np.zeros(max(len(q), 1), np.uint8)
    nf = lib().orc_hull_bruteforce(_p(q), len(q), _p(flags))
    return flags[:len(q)].astype(bool), nf


def hull_batch(pts, offsets, check_general=True, serial=False):
    """Per cloud: vertex flags (concatenated), face counts, FP64 volumes, general flags."""
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    offsets = np.ascontiguousarray(offsets, np.int32)
    n = len(offsets) - 1
    flags = np.zeros(max(len(pts), 1), np.uint8)
    nf = np.zeros(n, np.int32)
    vol = np.zeros(n, np.float64)
    gen = np.zeros(n, np.int32)
    if serial:
        bad = lib().orc_hull_batch_serial(_p(pts), _p(offsets), n, _p(flags), _p(nf), _p(vol))
    else:
        bad = lib().orc_hull_batch(_p(pts), _p(offsets), n, _p(flags), _p(nf), _p(vol), _p(gen),
                                   1 if check_general else 0)
    if bad:
        raise RuntimeError(f"oracle hull failed on {bad} clouds")
    return flags[:len(pts)].astype(bool), nf, vol, gen
