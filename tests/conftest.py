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
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from impeller_amd.runtime import Stream
    return Stream.current(0)
