import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def q():
    """The HIP library, initialised on device 0.  A GPU test fails (it does not
    skip) when the library or the device is missing."""
    import qkxtm_mpi_amd as qq
    qq.init(0)
    return qq
