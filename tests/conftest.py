import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP engine)")


def golden(name: str) -> str:
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def cube208():
    from tycho2_amd.mesh import read_smesh
    return read_smesh(golden("cube-208.smesh"))


@pytest.fixture(scope="session")
def cube12():
    from tycho2_amd.mesh import read_smesh
    return read_smesh(golden("cube-12.smesh"))


@pytest.fixture(scope="session")
def gold():
    return dict(np.load(golden("gold_cube208.npz")))


def gpu_available() -> bool:
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False
