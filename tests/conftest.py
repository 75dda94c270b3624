import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import ctypes

        from kd_slam2_amd import _lib

        h = ctypes.c_void_p()
        rc = _lib.lib().kdh_ctx_create(0, ctypes.byref(h))
        if rc == 0:
            _lib.lib().kdh_ctx_destroy(h)
            return True
    except Exception:
        return False
    return False


@pytest.fixture(scope="session")
def ctx():
    from kd_slam2_amd import Context

    c = Context(0)
    yield c
    c.close()
