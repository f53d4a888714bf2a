"""Shared pytest setup.

* ``gpu`` marks tests that need an AMD GPU; they call libmcgpu through its C
  ABI and compare with the CPU oracle.  Everything else runs on the CPU.
* The drop-in modules (core, models, utils, main, mcgpu) live under
  gpu-monte-carlo_amd/, which plays the role of the reference's project root.
"""

import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "gpu-monte-carlo_amd"
for p in (str(PKG), str(REPO)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running statistical check")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly (never skips) when the HIP library or device is missing."""
    import mcgpu
    from mcgpu import _lib
    _lib.load()
    _lib.ensure_device()
    return mcgpu


def ladder(R, lo, hi):
    return np.linspace(lo, hi, R)
