import os
import sys

import pytest

# Load PyTorch's HIP runtime before libbarbie_amd: the library then binds to
# the same libamdhip64.so.7 instead of bringing in a second runtime, and the
# tests that hand torch device buffers to the library share one runtime.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def pytest_runtest_logstart(nodeid, location):
    # Name the case before it touches the GPU, so a fault is attributable.
    print(f"\n[start] {nodeid}", flush=True)
