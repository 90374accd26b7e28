import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libraries")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def gpu():
    """The HIP device; GPU tests fail (not skip) when it is missing."""
    import torch

    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    from jaxphys_amd import _native

    _native.lib()
    return torch.device("cuda", 0)
