import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")


@pytest.fixture(scope="session")
def device():
    """One committed "hip" device for the GPU tests."""
    import openvkl_amd as vkl

    dev = vkl.Device("hip")
    yield dev
    dev.release()
