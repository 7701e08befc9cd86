import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def hip_device():
    """Skips when no HIP device is visible; the GPU tests are marked gpu."""
    import torch  # device presence only; the product path is the C ABI

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return 0
