import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through libharmonia_amd)")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is visible; make sure the HIP library is built and loads."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import harmonia_amd

    harmonia_amd.load()  # raises if libharmonia_amd.so is missing: no CPU fallback
    return torch.device("cuda:0")
