import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X); runs the HIP path through the C ABI")


@pytest.fixture(scope="session")
def dev():
    """The HIP device.  GPU tests fail (not skip) when the library is missing."""
    import torch
    import gcoospdm_amd
    gcoospdm_amd.lib()  # raises if libgcoospdm.so is not built
    assert torch.cuda.is_available(), "GPU test run without a HIP device"
    return torch.device("cuda:0")
