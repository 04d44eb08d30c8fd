import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device; fails loudly (no CPU fallback)."""
    from timemachine_amd import _lib

    _lib.load()
    n = _lib.device_count()
    assert n >= 1, "no HIP device visible: gpu tests need an MI355X"
    return _lib


@pytest.fixture()
def rng():
    return np.random.default_rng(2022)
