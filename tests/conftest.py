"""Shared fixtures. GPU tests carry @pytest.mark.gpu; everything else runs on CPU."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def orc():
    from oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def eb():
    """The HIP package (raises if libeuler_hip.so is missing: no fallback)."""
    from __graft_entry__ import load_package

    return load_package()


@pytest.fixture(scope="session")
def ex(eb):
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a HIP device")
    e = eb.ExecSpace(torch.device("cuda", 0))
    yield e
    e.close()


def random_prim(shape, seed, dtype=np.float64, d=(0.5, 2.0), p=(0.5, 2.0), sigma=0.3):
    """Seeded random primitive state (5, n2, n1, n0): rho, p uniform, u normal."""
    n0, n1, n2 = shape
    rng = np.random.default_rng(seed)
    out = np.empty((5, n2, n1, n0), dtype=np.float64)
    out[0] = rng.uniform(*d, size=(n2, n1, n0))
    out[1] = rng.uniform(*p, size=(n2, n1, n0))
    out[2:] = rng.normal(0.0, sigma, size=(3, n2, n1, n0))
    return out.astype(dtype)
