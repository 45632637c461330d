import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (str(REPO), str(REPO / "ram-scb_amd"), str(REPO / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libramgpu.so")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device count without initialising HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")
