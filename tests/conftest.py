import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libpdc.so")


@pytest.fixture(scope="session")
def gpu_available():
    from primal_dual_contacts_miror_amd import pdc
    try:
        return pdc.device_count() > 0
    except pdc.PdcError:
        return False


@pytest.fixture
def require_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("GPU test requested but no gfx950 device / libpdc.so is available")
