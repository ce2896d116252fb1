"""Every benchmark configuration exercised on the GPU against the oracle (or, where the oracle
cannot follow a long run, against FP64):

* C1 exactly: the 8^3 x 6 = 3,072-tet periodic vortex cube [0,10]^3, P1, Lax-Friedrichs,
  20 SSP-RK3 steps at CFL 0.3: the step-0 residual within 1e-12 and the step-20 state within
  1e-11 (norm-wise relative) of the oracle's;
* C2's protocol at full size: the 32^3 x 6 = 196,608-tet P3 Roe vortex cube; the GPU residual at
  step 0 and the GPU residual of the GPU's own step-200 state (200 steps at CFL 0.3) each compared
  element-wise with the oracle's residual of the same state (1e-12 of the residual's max norm);
* C5's scheme at reduced size: the compressible Taylor-Green vortex on [0,2pi]^3 (Mach 0.1), P4
  Roe on a 5^3 x 6 mesh, FP64, FP32 and BF16x3 contractions: each residual against the oracle
  (FP64 1e-12, mixed 2e-5), and 20 SSP-RK3 steps with the kinetic energy and enstrophy series
  (every 5 steps) of the mixed paths within 1e-8 of FP64's and the final states within 1e-6
  (relative L2).
C3 (mixed precision, jittered, P3) is covered by test_gpu_mixed.py at 16^3 x 6 and by the
benchmark at full size; C4 (the ensemble) by test_gpu_ensemble.py."""
import math

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX, rel_err
from fve_dg_amd import validate

pytestmark = pytest.mark.gpu

TGV = [1.4, 1.0, 1.0, 0.1]


def test_c1_exact(solver, oracle):
    print("[case] C1: 8^3x6 P1 lf, 20 steps", flush=True)
    mesh = F.Mesh.box(8, ub=(10, 10, 10), periodic=7)
    assert mesh.n_cells == 3072
    U0 = mesh.project(1, "vortex", VORTEX)
    solver.set_scheme(1, flux="lf").set_mesh(mesh)
    solver.upload(U0)
    m = mesh.export()
    assert rel_err(solver.residual(), oracle.residual(m, 1, U0, flux="lf")) < 1e-12
    dt = solver.stable_dt(0.3)
    solver.ssprk3(dt, 20)
    Ug = solver.download()
    Uo = oracle.ssprk3(m, 1, U0, dt, 20, flux="lf", threads=oracle.max_threads())
    err = rel_err(Ug, Uo)
    print(f"C1 state after 20 steps vs oracle: {err:.2e}", flush=True)
    assert err < 1e-11


def test_c2_protocol_full_size(solver, oracle):
    print("[case] C2: 32^3x6 P3 roe, residual at steps 0 and 200 against the oracle", flush=True)
    mesh = F.Mesh.box(32, ub=(10, 10, 10), periodic=7)
    assert mesh.n_cells == 196608
    U0 = mesh.project(3, "vortex", VORTEX)
    solver.set_scheme(3, flux="roe").set_mesh(mesh)
    solver.upload(U0)
    m = mesh.export()
    th = oracle.max_threads()
    R0 = solver.residual()
    e0 = rel_err(R0, oracle.residual(m, 3, U0, flux="roe", threads=th))
    dt = solver.stable_dt(0.3)
    solver.ssprk3(dt, 200)
    U200 = solver.download()
    assert np.isfinite(U200).all()
    R200 = solver.residual()
    e200 = rel_err(R200, oracle.residual(m, 3, U200, flux="roe", threads=th))
    print(f"C2 residual vs oracle: step 0 {e0:.2e}, step 200 {e200:.2e}", flush=True)
    assert e0 < 1e-12 and e200 < 1e-12


@pytest.fixture(scope="module")
def tgv_case():
    L = 2 * math.pi
    mesh = F.Mesh.box(5, ub=(L, L, L), periodic=7)
    return mesh, mesh.project(4, "tgv", TGV)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5), ("bf16x3", 2e-5)])
def test_c5_tgv_p4_residual(solver, oracle, tgv_case, prec, tol):
    """At Mach 0.1 the uniform pressure p0 = 1/(gamma M^2) = 71.4 enters the volume and face terms
    and cancels between them: the residual is ~60x smaller than its volume part, so both sum
    orders' roundings are measured against the volume part's max norm (the size of the terms)."""
    print(f"[case] C5 scheme: TGV P4 roe residual, {prec}", flush=True)
    mesh, U = tgv_case
    solver.set_scheme(4, flux="roe", precision=prec).set_mesh(mesh)
    solver.upload(U)
    m = mesh.export()
    Ro = oracle.residual(m, 4, U, flux="roe")
    scale = np.abs(oracle.residual(m, 4, U, flux="roe", parts=F.PART_CELLS)).max()
    err = np.abs(solver.residual() - Ro).max() / scale
    print(f"TGV P4 {prec} residual vs oracle: {err:.2e} of the volume part "
          f"(volume part / residual max norm: {scale / np.abs(Ro).max():.1f})", flush=True)
    assert err < tol
    assert prec == "fp64" or err > 1e-13  # the mixed contractions ran


def test_c5_tgv_p4_series():
    print("[case] C5 scheme: TGV P4 roe, 20 steps, fp64/fp32/bf16x3 series", flush=True)
    r = validate.run("tgv", n=5, order=4, steps=20, every=5, jitter=0.0, precisions=("fp64", "fp32", "bf16x3"))
    for prec in ("fp64", "fp32", "bf16x3"):
        S = np.array(r[prec]["series"])
        assert S.shape == (5, 7) and np.isfinite(S).all()
        assert S[-1, 1] <= S[0, 1]  # kinetic energy only dissipates
    for prec in ("fp32", "bf16x3"):
        dev = r[f"series_max_rel_deviation_{prec}"]
        print(f"TGV P4 {prec}: series deviation {dev}, state {r[f'l2_rel_deviation_{prec}_vs_fp64']:.2e}", flush=True)
        assert dev["kinetic_energy"] < 1e-8 and dev["enstrophy"] < 1e-8, prec
        assert r[f"l2_rel_deviation_{prec}_vs_fp64"] < 1e-6, prec
