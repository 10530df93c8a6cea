import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One engine context for the whole GPU session (fails loudly without a GPU)."""
    import torch

    import pinocchio_amd as pa

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()
    return pa.Context(0)
