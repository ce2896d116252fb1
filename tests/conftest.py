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


def project_pointwise(oracle, m, order, fn):
    """diagonal L2 projection of a pointwise state fn(x) -> (5,) with the oracle's basis and rule
    (exact for fields in the order's polynomial space)"""
    pts, w = oracle.quadrature(order, 0)
    og = oracle.geometry(m)
    Phi = np.array([oracle.basis(order, *x)[0] for x in pts])  # Q x N
    diag = (w[:, None] * Phi ** 2).sum(0)
    U = np.zeros((len(m["cells"]), 5 * Phi.shape[1]))
    for e, cell in enumerate(m["cells"]):
        X = m["points"][cell[0]] + pts @ og["A"][e]  # x(xi) = p0 + A^T xi, rows of A = edges
        vals = np.array([fn(x) for x in X])
        U[e] = ((w[:, None, None] * Phi[:, :, None] * vals[:, None, :]).sum(0) / diag[:, None]).ravel()
    return U
