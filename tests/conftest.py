import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU runs (reference scenario known answers); set HG_SLOW=1 to run")


def pytest_collection_modifyitems(config, items):
    # A ROCm GPU is visible when the kernel driver node exists.
    gpu_ok = os.path.exists("/dev/kfd")
    slow_ok = os.environ.get("HG_SLOW") == "1"
    for it in items:
        if "gpu" in it.keywords and not gpu_ok:
            it.add_marker(pytest.mark.skip(reason="no GPU visible"))
        if "slow" in it.keywords and not slow_ok:
            it.add_marker(pytest.mark.skip(reason="long CPU run; set HG_SLOW=1"))
