import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libctfp.so")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    # GPU tests are selected with -m gpu on a GPU machine; without a GPU they
    # fail instead of skipping so a missing device is never mistaken for a pass.
    if not gpu_available():
        pytest.fail("GPU test selected but no GPU is visible to torch")
