import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dg-sparsegrid_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLD = os.path.join(ROOT, "tests", "golden", "gold_fixtures.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libkmx.so")


@pytest.fixture(scope="session")
def gold():
    return np.load(GOLD)


@pytest.fixture(scope="session")
def kmx_gpu():
    """The product library on a live device; fails loudly if either is missing."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test run without a visible HIP device")
    import kmx

    kmx.version()  # raises if libkmx.so is missing
    return kmx
