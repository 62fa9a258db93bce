"""Test configuration: import paths and the `gpu` marker.

`-m "not gpu"` tests run anywhere (oracle vs golden vectors, host logic,
library loading); `-m gpu` tests need a gfx950 device and call the HIP path.
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "p2c-ccd_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 (MI355X) device")


def _gpu_present() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_present():
        pytest.skip("no GPU visible")
    import torch

    torch.cuda.init()
    return torch.device("cuda:0")
