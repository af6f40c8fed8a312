import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def hip_runtime():
    """Start torch's HIP runtime (when torch is installed) before
    libjurassic_gpu.so loads, so that the process uses one runtime for both
    (torch brings its own copy of the HIP runtime library)."""
    try:
        import torch
    except ImportError:
        return None
    if torch.cuda.is_available():
        torch.cuda.init()
    return torch


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O
