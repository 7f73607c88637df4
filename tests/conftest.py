import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmpc_hip.so)")


@pytest.fixture(scope="session")
def gpu_lib():
    """libmpc_hip.so with a visible GPU; the HIP path must be the one that runs."""
    from mpc_amd import lib
    L = lib()
    if L.mpc_device_count() < 1:
        pytest.fail("no GPU visible to libmpc_hip.so")
    return L
