import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "segment-tree-based-clipping_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsegclip.so")
    config.addinivalue_line("markers", "slow: full-size workload")


def gpu_available():
    try:
        import segclip
        import ctypes as C
        h = C.c_void_p()
        rc = segclip.lib().segclip_create(C.byref(h), 0, None)
        if rc == 0:
            segclip.lib().segclip_destroy(h)
            return True
    except Exception:
        return False
    return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device / libsegclip.so is usable")
    return True
