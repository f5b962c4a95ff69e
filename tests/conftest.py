import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libphysica_amd.so")


@pytest.fixture(scope="session")
def gpu_device():
    """The GPU the tests run on; fails (never falls back) if it is absent."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    from physica_amd import _capi

    _capi.library()  # raises if libphysica_amd.so is not built
    return torch.device("cuda:0")
