import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _gpu_available():
    # The AMD kernel driver node exists only where a GPU is visible.
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    """Skip GPU tests cleanly when no device is visible; never fall back."""
    if not _gpu_available():
        pytest.skip("no GPU visible")
    import annie_amd
    annie_amd.load_library()
    return True
