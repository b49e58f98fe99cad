import os
import sys

import pytest

# the live inflate queue's persistent kernel needs a hardware queue of its
# own: HIP maps streams onto GPU_MAX_HW_QUEUES of them (4 by default), and
# the test session creates more streams than that; set before HIP starts
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

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
