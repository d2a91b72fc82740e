import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "matrix-free-dealii-precice_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmfgpu.so on the device)")


@pytest.fixture(scope="session")
def cuda():
    """The HIP device; GPU tests fail loudly (no CPU fallback) when it is absent."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test run without a HIP device")
    import mfgpu
    mfgpu.lib()
    return torch.device("cuda:0")
