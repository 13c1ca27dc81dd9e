import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X / gfx950)")


@pytest.fixture(scope="session")
def stream():
    """The current HIP stream of device 0.  A GPU test that runs where no GPU
    is visible fails (it is deselected by -m "not gpu" on CPU-only hosts)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP GPU is visible", pytrace=False)
    from impeller_amd.runtime import Stream
    return Stream.current(0)
