import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "spmspv-adaptive_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the device)")


@pytest.fixture(scope="session")
def gpu():
    import spmspv_amd as sa
    if sa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path cannot run (no CPU fallback exists)")
    return sa
