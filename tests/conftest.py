import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import oracle_backend
    return oracle_backend()


@pytest.fixture(scope="session")
def hip():
    """The HIP library; on a GPU test this must load and see a device."""
    import ctypes as C

    import voxarmps_amd as vx
    be = vx.hip_backend()
    n = C.c_int(0)
    be.lib.vox_device_count(C.byref(n))
    if n.value < 1:
        raise RuntimeError("GPU test without a visible HIP device")
    return be
