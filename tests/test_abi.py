"""CPU tests of the drop-in boundary: the C-ABI library loads, exports every
function include/cuacd_amd.h declares, and the CPython extension exposes the
reference's `_gpu` method names (reference csrc/module.c:532-562)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "cuacd_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "return", "defined"}
    return sorted({n for n in names if n not in skip and (n.startswith("gpu_") or n == "lookahead_decompose")})


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "cuacd_amd", "libcuacd_hip.so"))
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_entry_points_are_declared():
    # every C function the reference extension calls (csrc/module.c) is part of the ABI
    ref = ["gpu_set_fatbin", "gpu_init", "gpu_destroy", "gpu_last_error", "gpu_heap_compact",
           "gpu_pool_usage", "gpu_heap_stats", "gpu_result_free", "gpu_test_warp_sort", "gpu_hull_dandc",
           "gpu_test_mesh_volume", "gpu_batch_mesh_volume", "gpu_kdop_hull", "gpu_test_hausdorff",
           "gpu_test_plane_cut", "gpu_mesh_audit", "gpu_test_postprocess_dc", "lookahead_decompose",
           "gpu_preprocess"]
    decl = set(_declared_functions())
    assert set(ref) <= decl


def test_extension_exposes_reference_methods():
    from cuacd_amd import _gpu
    for m in ["set_fatbin", "init", "destroy", "heap_compact", "pool_usage", "heap_stats", "test_warp_sort",
              "hull_dandc", "test_mesh_volume", "batch_mesh_volume", "test_plane_cut", "kdop_hull",
              "test_hausdorff", "test_postprocess_dc", "mesh_audit", "preprocess", "lookahead_decompose"]:
        assert hasattr(_gpu, m), m


def test_no_context_raises_runtime_error():
    from cuacd_amd import _gpu
    _gpu.destroy()
    with pytest.raises(RuntimeError, match="context not initialized"):
        _gpu.pool_usage()


def test_init_without_device_raises_cleanly():
    import cuacd_amd
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is visible; this checks the no-device error path")
    with pytest.raises(RuntimeError, match="cuacd error"):
        cuacd_amd.Context(device=0)


def test_kerr_decode_table_names():
    text = open(os.path.join(ROOT, "cuacd_amd", "csrc", "kerr.h")).read()
    bits = re.findall(r"#define (KERR_[A-Z0-9_]+)\s+\(1 << (\d+)\)", text)
    vals = [int(b) for _, b in bits]
    assert len(vals) == len(set(vals)), "duplicate KERR bits"
    for name, _ in bits:
        assert f"{{ {name}," in text, f"{name} missing from kerr_decode table"
