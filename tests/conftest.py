import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "2dmatching-cf4_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import ipm_oracle
    ipm_oracle.lib()
    return ipm_oracle


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
