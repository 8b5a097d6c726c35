import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a HIP device is visible.  The probe goes through
    libgmi.so itself, so a missing library is an error, not a skip."""
    from generic_mobility_interface_class_amd import capi
    if capi.device_count() < 1:
        pytest.skip("no HIP device visible")
    return 0
