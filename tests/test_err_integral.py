"""err_integral (fvdg_analytic_error; the per-variable relative L2 error of
examples/Vortex/VortexBatch3.cpp:349-366) pinned by a closed form.

Against a uniform field U_s = c, the state c + delta * phi_k e_v (mode k of variable v only) has
  err_v = |delta| sqrt(int phi_k^2) / (|c_v| sqrt(int 1)) = |delta| sqrt(6 mhat_k) / |c_v|
on any mesh, with the reference-element masses mhat_0 = int_T 1 = 1/6 and mhat_1 = int_T (4x - 1)^2
= 16/60 - 8/24 + 1/6 = 1/10 (monomial integrals a!b!c!/(a+b+c+3)!), and err = 0 for the other
variables.  A missing |det J|, a wrong weight sum, a missing square root or a swapped variable fails
it; the jittered periodic mesh makes the element volumes unequal."""
import numpy as np
import pytest

import fve_dg_amd as F

C0 = [1.2, 0.3, -0.4, 0.5, 2.7]


@pytest.mark.parametrize("order", [1, 3])
@pytest.mark.parametrize("k,mhat", [(0, 1 / 6), (1, 1 / 10)])
@pytest.mark.parametrize("v", [0, 2, 4])
def test_err_integral_closed_form(order, k, mhat, v):
    mesh = F.Mesh.box(3, ub=(2.0, 3.0, 1.5), periodic=7, jitter=0.15, seed=4)
    N = F.num_basis(order)
    U = mesh.project(order, "const", C0)
    delta = 0.0123
    U[:, 5 * k + v] += delta
    err = mesh.analytic_error(order, "const", C0, 0.0, U)
    expect = np.zeros(5)
    expect[v] = abs(delta) * np.sqrt(6 * mhat) / abs(C0[v])
    assert np.allclose(err, expect, rtol=1e-12, atol=1e-15), (err, expect)


def test_basis_mode_one_is_4x_minus_1():
    """the closed form above uses phi_1 = 4x - 1 on the unit tetrahedron"""
    for x, y, z in [(0.1, 0.2, 0.3), (0.7, 0.1, 0.05), (0.0, 0.0, 0.0)]:
        v, _ = F.basis(1, x, y, z)
        assert abs(v[1] - (4 * x - 1)) < 1e-14 and abs(v[0] - 1) < 1e-15
