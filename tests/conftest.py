import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, loaded; skips when no GPU is visible.  A missing
    library is an error, never a skip: there is no CPU fallback."""
    import compression_amd as mdz
    mdz.lib()
    if mdz.device_count() < 1:
        pytest.skip("no GPU visible")
    return mdz
