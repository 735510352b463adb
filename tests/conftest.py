import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libht3d.so")


@pytest.fixture(scope="session")
def ht():
    """The HIP package, initialised on device 0.  Fails loudly if libht3d.so is
    missing or no GPU is visible: GPU tests never fall back to the CPU."""
    import hydrothermal3d_amd as ht3d
    ht3d._lib.ensure_init()
    return ht3d


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


def farr(shape, rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, size=shape))
