import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# GPU tests of several minutes each (whole tests at the configurations'
# sizes) run only when this is set to 1, so that the default GPU suite stays
# within a quarter of an hour.
LONG_ENV = "ANNIE_LONG_TESTS"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running CPU test")
    config.addinivalue_line("markers", f"long: GPU test of minutes; runs when {LONG_ENV}=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get(LONG_ENV) == "1":
        return
    skip = pytest.mark.skip(reason=f"long GPU test (minutes): set {LONG_ENV}=1 to run it")
    for item in items:
        if "long" in item.keywords:
            item.add_marker(skip)


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
