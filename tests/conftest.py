import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the gfx950 library)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The gfx950 library; a missing library is a failure, not a skip."""
    import ifos_amd
    L = ifos_amd.lib()
    if L.ifos_device_count() < 1:
        pytest.skip("no GPU visible")
    return L
