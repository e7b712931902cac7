import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def gpu_available():
    try:
        import tfhe_amd
        return tfhe_amd.lib().tfhe_amd_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    import tfhe_amd
    if tfhe_amd.lib().tfhe_amd_device_count() < 1:
        pytest.fail("GPU test collected but no GPU is visible to libtfhe_amd.so")
    return tfhe_amd
