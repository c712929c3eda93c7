import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def klrtm():
    """The product library; GPU tests fail loudly when it is not built."""
    from konglong_amd import _lib
    _lib.lib()
    if _lib.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return _lib
