import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "rust-la_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx(torch_cuda):
    import la_hip
    c = la_hip.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch


@pytest.fixture(scope="module")
def tctx(torch_cuda):
    """A context whose stream is torch's current stream, so library work and
    torch tensor operations are ordered without extra synchronisation."""
    import la_hip
    c = la_hip.Context(0)
    c.use_torch_stream()
    yield c
    c.close()
