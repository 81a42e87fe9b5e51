import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "amrex-euler-allmach-explicit-imex_amd"
for p in (ROOT / "oracle", PKG, ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the C-ABI)")


@pytest.fixture(scope="session")
def ecor():
    import ecor as mod
    mod.build()
    return mod


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
