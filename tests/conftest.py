import os
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqed3 on the GPU)")


def has_gpu():
    try:
        import qed3_amd._lib as L
        lib = L.load()
        import ctypes as C
        p = C.c_void_p()
        rc = lib.qed3_malloc(C.byref(p), 16)
        if rc == 0:
            lib.qed3_free(p)
            return True
    except Exception:
        return False
    return False


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.fail("GPU test selected but no usable HIP device / libqed3.so")
    return True


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def at_default(mean_ell):
    """a_t = <ell>/2: satisfies the reference's sqrt3<ell>/a_t >= 4/sqrt3 (hmc.cu:169)."""
    return 0.5 * mean_ell
