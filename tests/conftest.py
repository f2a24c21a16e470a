import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libmcpgpu.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def handle():
    """One libmcpgpu handle for the GPU test session; fails loudly when the library is missing."""
    import mcpnet_amd
    h = mcpnet_amd.Handle(int(os.environ.get("MCPG_DEVICE", "0")))
    yield h
    h.close()


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible ROCm device"
    return torch
