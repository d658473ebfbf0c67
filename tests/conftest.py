import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP product path)")


@pytest.fixture(scope="session")
def torch_cuda():
    """torch with a visible GPU; GPU tests fail (not skip) without one."""
    import torch
    assert torch.cuda.is_available(), "GPU test started without a visible GPU"
    return torch


@pytest.fixture(scope="session")
def product():
    """The HIP product library; fails loudly if it was not built."""
    import exadg_amd
    exadg_amd.lib()
    return exadg_amd
