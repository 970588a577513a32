import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole GPU test session (HIP path only)."""
    import cuacd_amd
    c = cuacd_amd.Context(device=0)
    yield c
    c.close()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def load_obj_text(text):
    """Minimal OBJ reader (v / f lines, 1-based indices, polygons fanned)."""
    vs, fs = [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            vs.append([float(x) for x in parts[1:4]])
        elif parts[0] == "f":
            idx = [int(p.split("/")[0]) - 1 for p in parts[1:]]
            for k in range(1, len(idx) - 1):
                fs.append([idx[0], idx[k], idx[k + 1]])
    return np.array(vs, np.float32), np.array(fs, np.int32)
