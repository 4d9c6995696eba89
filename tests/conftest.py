import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libqms.so kernels)")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to count devices; not used for compute)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")
