import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def nbl():
    """libnbody.so bindings (host entry points work without a GPU)."""
    from nbodysimulation_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: GPU tests are only collected with -m gpu on a GPU box;
    if the device is missing there, fail loudly instead of falling back."""
    from nbodysimulation_amd import device_info
    info = device_info()
    if info is None:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    return info
