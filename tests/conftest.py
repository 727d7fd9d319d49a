"""pytest configuration: the `gpu` marker, import paths, shared fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "wgpu-spectral-rt_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libwsrt.so")


@pytest.fixture(scope="session")
def device():
    import wsrt

    if wsrt.device_count() == 0:
        pytest.fail("GPU test run without a HIP device")
    return wsrt.Device(0)


@pytest.fixture(scope="session")
def cornell_pair(device):
    """The Cornell box (cornell_box.rs) built by the HIP path and by the oracle."""
    import oracle as O
    import wsrt

    return wsrt.create_cornell_box(device), O.cornell_box()
