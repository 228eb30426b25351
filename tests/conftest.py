import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def device():
    import luisacompute_amd as lc
    # No skip: a GPU test that runs without the native library or a GPU must fail loudly.
    return lc.Device(0)


@pytest.fixture(scope="session")
def stream(device):
    return device.create_stream()
