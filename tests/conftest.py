import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bwa-mem4_amd")
for p in (PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_present():
    try:
        import bwg
        bwg.library()
    except Exception:
        return False
    import ctypes
    h = ctypes.c_void_p()
    rc = bwg.library().bwg_create(0, ctypes.byref(h))
    if rc == 0:
        bwg.library().bwg_destroy(h)
        return True
    return False


@pytest.fixture(scope="session")
def gpu_ctx():
    """A libbwg context on device 0.  GPU tests fail loudly (never skip) when it cannot be made."""
    import bwg
    ctx = bwg.Context(0)
    yield ctx
    ctx.close()
