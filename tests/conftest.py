"""Shared pytest configuration.

``-m gpu`` tests need an AMD GPU and call libmatvis_hip.so through its C ABI;
``-m "not gpu"`` tests run on the CPU (oracle, host logic, ABI surface).
"""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(20261202)
