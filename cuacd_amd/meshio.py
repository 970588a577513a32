"""Mesh files in and out without third-party mesh libraries.

* ``read_stl`` - binary or ASCII STL; corners with identical float32
  coordinates become one vertex (numbered in order of first occurrence), so
  a closed STL surface comes back as an indexed watertight mesh.
* ``read_obj`` / ``write_obj`` - Wavefront OBJ (``v`` and ``f`` records;
  polygons are fanned from their first corner; texture and normal indices are
  ignored).
* ``read_mesh`` - either, by file extension.

Arrays are float32 vertices [n, 3] and int32 triangles [m, 3].
"""
import os

import numpy as np


def _merge_corners(corners):
    """Indexed mesh from per-triangle corner coordinates [m, 3, 3]."""
    c = np.ascontiguousarray(corners, np.float32).reshape(-1, 3)
    if len(c) == 0:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32)
    key = c.view(np.dtype((np.void, 12))).ravel()
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # renumber by first occurrence so the order follows the file
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(order), np.int64)
    rank[order] = np.arange(len(order))
    verts = c[first[order]]
    tris = rank[inv.ravel()].reshape(-1, 3).astype(np.int32)
    return verts, tris


def read_stl(path):
    data = open(path, "rb").read()
    if len(data) >= 84:
        n = int(np.frombuffer(data, np.uint32, 1, 80)[0])
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]), n, 84)
            return _merge_corners(rec["v"])
    text = data.decode("ascii", errors="replace")
    if not text.lstrip().startswith("solid"):
        raise ValueError(f"{path}: neither a binary STL (size mismatch) nor an ASCII one")
    vs = [[float(x) for x in line.split()[1:4]] for line in text.splitlines() if line.strip().startswith("vertex")]
    if len(vs) % 3:
        raise ValueError(f"{path}: ASCII STL with {len(vs)} vertex records (not a multiple of 3)")
    return _merge_corners(np.array(vs, np.float32).reshape(-1, 3, 3))


def read_obj(path):
    vs, fs = [], []
    with open(path, "r", errors="replace") as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                vs.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                idx = []
                for q in p[1:]:
                    k = int(q.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(vs) + k)
                for j in range(1, len(idx) - 1):
                    fs.append([idx[0], idx[j], idx[j + 1]])
    v = np.array(vs, np.float32).reshape(-1, 3)
    t = np.array(fs, np.int32).reshape(-1, 3)
    if len(t) and (t.min() < 0 or t.max() >= len(v)):
        raise ValueError(f"{path}: face index out of range")
    return v, t


def write_obj(path, verts, tris):
    verts = np.asarray(verts, np.float64).reshape(-1, 3)
    tris = np.asarray(tris, np.int64).reshape(-1, 3)
    with open(path, "w") as f:
        for x, y, z in verts:
            f.write(f"v {x:.9g} {y:.9g} {z:.9g}\n")
        for a, b, c in tris + 1:
            f.write(f"f {a} {b} {c}\n")


def read_mesh(path):
    ext = os.path.splitext(path)[1].lower()
    if ext == ".stl":
        return read_stl(path)
    if ext == ".obj":
        return read_obj(path)
    raise ValueError(f"{path}: unsupported mesh format {ext!r} (STL and OBJ are read)")
