import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; parity tests through the C ABI")


@pytest.fixture(scope="session")
def gpu():
    """Every GPU test requires a visible device and the native library."""
    import padenti_amd
    n = padenti_amd.device_count()
    assert n > 0, "no HIP device visible: GPU tests must run on a GPU"
    return 0
