import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU runs (reference scenario known answers)")


def pytest_collection_modifyitems(config, items):
    gpu_ok = None
    for it in items:
        if "gpu" in it.keywords:
            if gpu_ok is None:
                try:
                    import torch  # noqa: F401
                    gpu_ok = torch.cuda.is_available()
                except Exception:
                    gpu_ok = False
            if not gpu_ok:
                it.add_marker(pytest.mark.skip(reason="no GPU visible"))
