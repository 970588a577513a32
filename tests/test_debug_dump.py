"""The search's debug output: the part-dump reader on files in the
reference's layout (CPU), and on the GPU a decomposition with
CUACD_DUMP_PARTS_DIR and verbose=1 whose dumps and tables match what the
search returns (reference docs/debugging.md, csrc/lookahead.c:373-436)."""
import os
import struct

import numpy as np
import pytest

from cuacd_amd import datasets, debug


def _write(path, v, t, mv, hv):
    with open(path, "wb") as f:
        f.write(struct.pack("<2i2f", len(v), len(t), mv, hv))
        f.write(np.asarray(v, "<f4").tobytes())
        f.write(np.asarray(t, "<i4").tobytes())


def test_reader_round_trip(tmp_path):
    v, t = datasets.l_shape_union()
    _write(tmp_path / "iter000_part000.bin", v, t, 3.0, 3.5)
    _write(tmp_path / "iter001_part002.bin", v[:5], t[:1], 0.25, 0.5)
    _write(tmp_path / "mesh004_iter002_part001.bin", v, t, 1.0, 2.0)
    (tmp_path / "notes.txt").write_text("not a dump")
    d = debug.load_dump_dir(str(tmp_path))
    assert sorted(d) == [(0, 0, 0), (0, 1, 2), (4, 2, 1)]
    rv, rt, mv, hv = d[(0, 0, 0)]
    assert rv.tobytes() == v.astype("<f4").tobytes() and np.array_equal(rt, t) and (mv, hv) == (3.0, 3.5)
    assert d[(0, 1, 2)][0].shape == (5, 3) and d[(0, 1, 2)][1].shape == (1, 3)


def test_reader_rejects_truncated_files(tmp_path):
    v, t = datasets.l_shape_union()
    p = tmp_path / "iter000_part000.bin"
    _write(p, v, t, 3.0, 3.5)
    p.write_bytes(p.read_bytes()[:-4])
    with pytest.raises(ValueError):
        debug.read_part_dump(str(p))


@pytest.mark.gpu
def test_dump_and_verbose_tables_of_a_search(ctx, tmp_path, monkeypatch, capfd):
    import oracle
    v, t = datasets.l_shape_union()
    monkeypatch.setenv("CUACD_DUMP_PARTS_DIR", str(tmp_path))
    parts = ctx.lookahead_decompose(v, t, max_iters=20, width=12, width2=3, n_concave_edges=4, verbose=1)
    err = capfd.readouterr().err
    d = debug.load_dump_dir(str(tmp_path))
    iters = sorted({k[1] for k in d})
    assert iters and iters[0] == 0
    # iteration 0 holds the input mesh, its FP64 volume rounded to float
    (v0, t0, mv0, hv0), = [x for k, x in d.items() if k[1] == 0]
    assert v0.tobytes() == v.tobytes() and np.array_equal(t0, t)
    assert abs(mv0 - 3.0) < 1e-6 and abs(hv0 - 3.5) < 1e-6
    # every dumped part is a closed mesh with its recorded volume
    for (_, it, p), (pv, pt, mv, hv) in d.items():
        assert oracle.edge_manifold_report(pt) == (0, 0)
        assert abs(abs(oracle.mesh_volume(pv, pt)) - mv) <= 1e-5 * max(mv, 1e-9)
    # the last dumped iteration's parts are the final parts (the search
    # stops at the iteration that finds nothing to cut)
    last = sorted((x for k, x in d.items() if k[1] == iters[-1]), key=lambda x: x[0].tobytes())
    final = sorted(parts, key=lambda p: p[0].tobytes())
    assert len(last) == len(final)
    for (pv, pt, _, _), (fv, ft, _, _) in zip(last, final):
        assert pv.tobytes() == fv.tobytes() and np.array_equal(pt, ft)
    # verbose: one table per iteration with the reference's columns and
    # n_cutting lines
    assert err.count("part   nv   nt  mesh_vol  hull_vol    rv_cost  hausdorff  full_cost") == len(iters)
    assert "[la] iter 0: n_cutting=1" in err
