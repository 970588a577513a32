"""Readers for the search's debug output.

``CUACD_DUMP_PARTS_DIR=<dir>`` makes every lookahead iteration write each
part of the decomposition, after the iteration's cost sort, to
``<dir>/iter%03d_part%03d.bin`` (``mesh%03d_`` prefixed when a batched call
has several meshes), in the reference's layout (reference
docs/debugging.md, csrc/lookahead.c:401-431): int32 ``{nv, nt}``, float32
``{mesh_vol, hull_vol}``, float32 vertices ``[nv, 3]``, int32 triangles
``[nt, 3]``, little-endian.
"""
import os
import re

import numpy as np

_NAME = re.compile(r"^(?:mesh(\d{3,})_)?iter(\d{3,})_part(\d{3,})\.bin$")


def read_part_dump(path):
    """(verts float32 [nv, 3], tris int32 [nt, 3], mesh_vol, hull_vol) of one dump file."""
    data = open(path, "rb").read()
    if len(data) < 16:
        raise ValueError(f"{path}: {len(data)} bytes, shorter than the 16-byte header")
    nv, nt = (int(x) for x in np.frombuffer(data, "<i4", 2, 0))
    mv, hv = (float(x) for x in np.frombuffer(data, "<f4", 2, 8))
    if nv < 0 or nt < 0 or len(data) != 16 + 12 * nv + 12 * nt:
        raise ValueError(f"{path}: header says {nv} vertices and {nt} triangles, file has {len(data)} bytes")
    v = np.frombuffer(data, "<f4", 3 * nv, 16).reshape(nv, 3).copy()
    t = np.frombuffer(data, "<i4", 3 * nt, 16 + 12 * nv).reshape(nt, 3).copy()
    return v, t, mv, hv


def load_dump_dir(path):
    """{(mesh, iteration, part): (verts, tris, mesh_vol, hull_vol)} of every
    dump file in a directory (mesh is 0 for single-mesh calls)."""
    out = {}
    for name in sorted(os.listdir(path)):
        m = _NAME.match(name)
        if not m:
            continue
        key = (int(m.group(1) or 0), int(m.group(2)), int(m.group(3)))
        out[key] = read_part_dump(os.path.join(path, name))
    return out
