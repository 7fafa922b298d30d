import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """libgas.so loaded with a visible GPU; fails (not skips) otherwise."""
    import vertexapi2_amd as gas
    n = gas.device_count()
    assert n > 0, "no HIP device visible: GPU tests need the MI355X"
    return gas
