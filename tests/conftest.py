import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpbf_amd.so")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; GPU tests fail loudly when it or the device is missing."""
    from pbf_amd import _lib
    lib = _lib.load()
    if lib.pbf_device_count() < 1:
        pytest.fail("no HIP device visible")
    return lib
