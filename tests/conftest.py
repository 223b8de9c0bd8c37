import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on device 0)")


@pytest.fixture(scope="session")
def gpu_ready():
    """Fails (never skips) when no device is visible: a GPU test must run the HIP path."""
    from pscbfm_amd import device_count
    n = device_count()
    assert n > 0, "no HIP device visible to a test marked gpu"
    return n
