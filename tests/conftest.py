"""Shared fixtures and markers.

``gpu`` marks tests that need an MI355X and the built ``libhfjk.so``; the
CPU suite (``-m "not gpu"``) covers the oracle, the generator, the layout,
the host logic and the exported C ABI.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def water():
    """Stored water/STO-3G fixtures (PySCF 2.14 values held by the reference)."""
    data = np.load(os.path.join(GOLDEN, "water_sto3g.npz"))
    return {k: data[k] for k in data.files}


@pytest.fixture(scope="session")
def synthetic_c0():
    data = np.load(os.path.join(GOLDEN, "synthetic_c0.npz"))
    return {k: data[k] for k in data.files}
