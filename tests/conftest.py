import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built native library")


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a gpu-marked test that runs without a device must fail loudly."""
    import palabos_amd.native as native
    native.init(0)
    return True
