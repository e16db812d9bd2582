import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built engine")
    config.addinivalue_line("markers", "slow: large sizes, minutes of CPU time")


def _has_gpu():
    try:
        from ezmock_amd import device_count
        return device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_available():
    if not _has_gpu():
        pytest.fail("GPU test selected but no HIP device / engine library is available")
    return True


@pytest.fixture(scope="session")
def pk_table():
    from ezmock_amd import inputs
    return inputs.eh_nowiggle_pk()
