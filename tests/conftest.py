import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libftmi.so")
    config.addinivalue_line("markers", "slow: a long oracle run (minutes of NumPy on the CPU); runs when FTMI_SLOW=1")


def _gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if not os.environ.get("FTMI_SLOW"):
        slow = pytest.mark.skip(reason="long oracle run: set FTMI_SLOW=1 to run it")
        for item in items:
            if "slow" in item.keywords:
                item.add_marker(slow)
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no HIP device in this process")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
