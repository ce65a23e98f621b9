import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def has_gpu():
    try:
        import fornos_amd
        return fornos_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    # GPU tests run only where a device is expected: a missing device there
    # is a failure, not a reason to skip.
    if not has_gpu():
        pytest.fail("no HIP device visible (fornos_device_count() == 0)")
    return True
