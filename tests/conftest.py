import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available():
    try:
        import xkblas_amd
        return xkblas_amd.lib().xkblas_get_devicecount() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def xk():
    import xkblas_amd
    if not gpu_available():
        pytest.skip("no GPU visible")
    xkblas_amd.lib().xkblas_init()
    yield xkblas_amd
    xkblas_amd.invalidate()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
