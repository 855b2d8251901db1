"""Test configuration: import paths and the ``gpu`` marker.

Tests marked ``gpu`` need an MI355X (gfx950) and the built libzolo.so; the
rest run on any CPU host.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "zolo-svd_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the HIP library")


@pytest.fixture(scope="session")
def zl():
    import zolo
    return zolo


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no GPU is visible")
    return torch
