import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "nemo-fmi-devel_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def hip():
    """The HIP library in device-pointer mode.  Fails (never skips) when the
    library or the GPU is missing: a GPU test must exercise the native path."""
    import torch
    import nemo_hip
    assert torch.cuda.is_available(), "gpu tests need a GPU"
    h = nemo_hip.get(0, nemo_hip.NEMO_HIP_PTR_DEVICE)
    return h
