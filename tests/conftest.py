from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and the built libtopohip.so")


@pytest.fixture
def golden():
    def load(name: str) -> dict:
        with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    return load


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests must run on a GPU box; fail loudly otherwise."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch.cuda.is_available() is False")
    from topogpu_amd.hip import lib

    lib.load()
    return torch.device("cuda", 0)
