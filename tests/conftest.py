import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "anisotropic-ml_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libnepx.so")


@pytest.fixture(scope="session")
def gpu():
    """The HIP path must be what runs: no silent fallback when this fixture is used."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    import nepx  # noqa: F401  (raises when libnepx.so is missing)
    return torch.device("cuda:0")
