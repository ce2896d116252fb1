"""Shared test setup: repository root on sys.path, the ``gpu`` marker, and
fixtures for the library binding, the CPU oracle and the reference build."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import fve_dg_amd as F  # noqa: E402
from oracle.binding import Oracle, Reference  # noqa: E402

# vortex of the benchmark configuration: gamma, strength, centre (x0, y0), free stream (u, v), periods
VORTEX = [1.4, 5.0, 5.0, 5.0, 1.0, 1.0, 10.0, 10.0]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    if not Reference.available():
        pytest.skip("reference build oracle/_ref/libref.so not present (needs the reference tree)")
    return Reference()


@pytest.fixture(scope="session")
def solver():
    s = F.Solver(0)
    yield s
    s.close()


def rel_err(a, b):
    """max |a-b| / max |b| (norm-wise relative deviation)"""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
