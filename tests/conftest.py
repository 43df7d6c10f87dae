import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libisp_bm3d.so")


@pytest.fixture(scope="session")
def gpu():
    """Skips when no HIP device is present; otherwise requires the HIP library."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from isp_mw_amd import _lib
    _lib.lib()  # raises if libisp_bm3d.so is missing: no fallback
    return torch.device("cuda:0")
