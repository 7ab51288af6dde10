import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")


def gpu_available():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
