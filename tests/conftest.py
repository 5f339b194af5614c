import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available():
    try:
        import torch  # noqa: F401  (only to count devices)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")
