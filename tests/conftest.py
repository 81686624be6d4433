import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def hip():
    """The HIP C-ABI library bound through ctypes (GPU tests only)."""
    from masterthesis_amd import capi
    lib = capi.load()
    if capi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return lib
