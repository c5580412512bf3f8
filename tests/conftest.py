import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when the HIP library or the device is unusable."""
    import luminara_amd as lm
    info = lm.device_info()
    print(f"[gpu] {info}", flush=True)
    return info
