"""Failure detection of the explicit stepper, forced from the host (no kernel faults):

* a NaN in one element's coefficients makes fvdg_max_wave_speed NaN and fvdg_stable_dt return
  FVDG_ERR_NONFINITE (the NaN-keeping maximum; fmax would drop it);
* an SSP-RK3 step from that state returns FVDG_ERR_NONFINITE in every precision;
* a negative mean density in one element with finite values everywhere (tiny dt, so the step
  keeps it negative) returns FVDG_ERR_NONFINITE through the mean-density check;
* a healthy state keeps returning OK after a failed call (the flag is cleared per call).
The reference has no such check: its drivers stop when the time step is not positive
(examples/Sedov/SedovBatch.cpp:200)."""
import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=3)
    order = 3
    return mesh, order, mesh.project(order, "vortex", VORTEX)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_nan_state_is_reported(solver, case, precision):
    mesh, order, U = case
    solver.set_scheme(order, flux="roe", precision=precision).set_mesh(mesh)
    bad = U.copy()
    bad[17, 5 * 3 + 4] = np.nan  # energy coefficient of mode 3 in one element
    solver.upload(bad)
    assert np.isnan(solver.max_wave_speed())
    with pytest.raises(F.FvdgError) as ei:
        solver.stable_dt(0.3)
    assert ei.value.status == F.FVDG_ERR_NONFINITE
    with pytest.raises(F.FvdgError) as ei:
        solver.ssprk3(1e-3, 1)
    assert ei.value.status == F.FVDG_ERR_NONFINITE
    # a healthy state afterwards: OK again
    solver.upload(U)
    dt = solver.stable_dt(0.3)
    solver.ssprk3(dt, 2)
    assert np.isfinite(solver.download()).all()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_negative_mean_density_is_reported(solver, case, precision):
    mesh, order, _ = case
    N = F.num_basis(order)
    U = np.zeros((mesh.n_cells, 5 * N))
    U[:, 0] = 1.0
    U[:, 4] = 2.5  # rest state, p = 1
    U[5, 0] = -1e-3  # one element with a negative mean density, finite everywhere
    solver.set_scheme(order, flux="lf", precision=precision).set_mesh(mesh)
    solver.upload(U)
    with pytest.raises(F.FvdgError) as ei:
        solver.ssprk3_stage(1e-12, 0)
    assert ei.value.status == F.FVDG_ERR_NONFINITE
    out = solver.download()
    assert np.isfinite(out).all()
