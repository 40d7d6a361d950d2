import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def hip_device():
    """Initialise the HIP library on device 0; GPU tests fail loudly without it."""
    from iwf_mfree_gpu_3d_amd import _lib
    _lib.init(0)
    return 0
