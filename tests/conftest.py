import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libslgpu.so")
    config.addinivalue_line("markers", "slow: longer statistical runs")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return dict(np.load(os.path.join(GOLDEN, name)))
    return load


@pytest.fixture(scope="session")
def c1():
    from superlite_amd import synth
    return synth.build("c1")


@pytest.fixture(scope="session")
def tables():
    from superlite_amd import synth
    return synth.load_tables()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
