import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP path")


@pytest.fixture(scope="session")
def ctx():
    """A pf_ctx on device 0.  The HIP library must be built: a missing library
    is an error, not a skip.  Only the absence of a device skips."""
    from periodfind_amd import capi

    capi.lib()  # raises ImportError if libpfhip.so is missing
    if capi.device_count() == 0:
        pytest.skip("no HIP device visible")
    c = capi.Context(0)
    yield c
    c.close()
