import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def sim():
    """A HIP simulator; fails loudly (never falls back) when the library or
    the GPU is missing."""
    from openqarp_amd.simulator import QarpSimulator
    return QarpSimulator()


@pytest.fixture(scope="session")
def sim64():
    from openqarp_amd.simulator import QarpSimulator
    return QarpSimulator(precision="complex64")
