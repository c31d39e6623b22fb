import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


@pytest.fixture(autouse=True)
def _fresh_configuration():
    from pyop2_amd.configuration import configuration
    saved = dict(configuration)
    yield
    configuration.unsafe_reconfigure(**saved)


@pytest.fixture(scope="session")
def device():
    from pyop2_amd import device as dev
    d = dev.device()
    print(f"[device] {d} {dev.torch().cuda.get_device_name(d)}", flush=True)
    return d
