"""GPU mixed precision: FVDG_PREC_FP32 (volume and face contractions in FP32 with FP32
accumulation) and FVDG_PREC_BF16X3 (the volume contractions on the bf16 matrix cores with
three-part bf16 operands and FP32 accumulation, the face contractions as FP32); pointwise/Riemann
fluxes, stage update and state in FP64 in both.

Bars (written per test): the mixed residual is within 2e-5 (norm-wise relative) of the FP64
oracle residual and differs from it (the mixed kernels really ran); a trajectory stays within
1e-6 relative L2 of the FP64 trajectory; its error against the analytic vortex equals the FP64
path's to 1e-3 relative; the cell-mean integrals stay conserved.
"""
import os

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX, rel_err

pytestmark = pytest.mark.gpu


def l2_rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


MIXED = ["fp32", "bf16x3"]


@pytest.mark.parametrize("prec", MIXED)
@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("flux", ["lf", "roe"])
def test_fp32_residual_close_to_oracle(solver, oracle, order, flux, prec):
    print(f"[case] {prec} residual order={order} flux={flux}", flush=True)
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=21)
    U = mesh.project(order, "vortex", VORTEX)
    U *= 1 + 1e-2 * np.random.default_rng(order).standard_normal(U.shape)
    solver.set_scheme(order, flux=flux, precision=prec).set_mesh(mesh)
    solver.upload(U)
    Rm = solver.residual()
    Ro = oracle.residual(mesh.export(), order, U, flux=flux)
    err = rel_err(Rm, Ro)
    assert err < 2e-5
    assert order == 0 or err > 1e-12  # FP32 contractions were used


@pytest.mark.parametrize("prec", MIXED)
@pytest.mark.parametrize("order", [2, 3, 4])
def test_fp32_parts_and_boundaries(solver, oracle, prec, order):
    """Dirichlet (x/y) and Pseudo3DZ (z) faces in the mixed path at P2-P4 (the orders whose
    face-flux pass holds the coefficients in registers), each residual part against the oracle."""
    print(f"[case] {prec} boundaries P{order}", flush=True)
    mesh = F.Mesh.box((4, 4, 2), ub=(10, 10, 2.5), split=5)
    mesh.tag_pseudo3dz()
    bc = mesh.boundary_types()
    params = [1.4, 5.0, 4.0, 4.5, 1.0, 0.5, 0.0, 0.0]
    U = mesh.project(order, "vortex", params)
    solver.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
    solver.upload(U)
    _, xyz = solver.dirichlet_points()
    UD = np.array([[oracle.vortex_state(params, 0.0, x) for x in row] for row in xyz])
    solver.set_dirichlet_states(UD)
    m = mesh.export()
    for parts in (1, 2, 4, 7):
        Rm = solver.residual(parts=parts)
        Ro = oracle.residual(m, order, U, flux="roe", bc=bc, UD=UD, parts=parts)
        assert rel_err(Rm, Ro) < 2e-5, parts


def test_precision_switch_keeps_the_state(solver):
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7)
    U = mesh.project(2, "vortex", VORTEX)
    solver.set_scheme(2, precision="fp64").set_mesh(mesh)
    solver.upload(U)
    solver.set_scheme(2, precision="fp32")
    assert np.array_equal(solver.download(), U)
    r32 = solver.residual()
    solver.set_scheme(2, precision="fp64")
    r64 = solver.residual()
    assert 0 < rel_err(r32, r64) < 2e-5


@pytest.mark.parametrize("order", [2, 3, 4])
def test_bf16x3_stage_on_face_tiles(solver, oracle, order):
    """BF16x3 at P2-P4: the stage runs the face tiles and the element kernel with every contraction
    on the bf16 matrix cores (k_face_tile_bf3<P>, k_elem_mfma<P,float,true,true,true>); its
    residual differs from the FP32 path's in the last bits and stays within 2e-5 of the oracle."""
    print(f"[case] bf16x3 face-tile stage P{order}", flush=True)
    mesh = F.Mesh.box(4, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=9)
    U = mesh.project(order, "vortex", VORTEX)
    U *= 1 + 1e-2 * np.random.default_rng(order).standard_normal(U.shape)
    R = {}
    for p in ("fp32", "bf16x3"):
        solver.set_scheme(order, flux="roe", precision=p).set_mesh(mesh)
        k = solver.stage_kernels()
        flux = f"k_face_tile_bf3<{order}>" if p == "bf16x3" else f"k_face_tile<{order},float>"
        assert k["volume"] == "" and k["flux"] == flux, k
        assert k["test"] == f"k_elem_mfma<{order},float,true,true" + (",true>" if p == "bf16x3" else ">"), k
        solver.upload(U)
        R[p] = solver.residual()
    assert not np.array_equal(R["bf16x3"], R["fp32"])
    Ro = oracle.residual(mesh.export(), order, U, flux="roe")
    for p in ("fp32", "bf16x3"):
        err = rel_err(R[p], Ro)
        print(f"{p} residual vs oracle: {err:.3e}", flush=True)
        assert err < 2e-5, p


@pytest.mark.parametrize("order", [1, 3, 5])
def test_bf16x3_volume_is_its_own_kernel(solver, order):
    """The BF16x3 volume term differs from the FP32 one in the last bits (another contraction
    ran) and both stay within 2e-5 of the FP64 volume term, relative to its norm."""
    print(f"[case] bf16x3 volume order={order}", flush=True)
    mesh = F.Mesh.box(4, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=5)
    U = mesh.project(order, "vortex", VORTEX)
    U *= 1 + 1e-2 * np.random.default_rng(order).standard_normal(U.shape)
    R = {}
    for p in ("fp64", "fp32", "bf16x3"):
        solver.set_scheme(order, flux="roe", precision=p).set_mesh(mesh)
        solver.upload(U)
        R[p] = solver.residual(parts=F.PART_CELLS)
    assert not np.array_equal(R["bf16x3"], R["fp32"])
    for p in ("fp32", "bf16x3"):
        err = rel_err(R[p], R["fp64"])
        print(f"{p} volume term vs fp64: {err:.3e}", flush=True)
        assert 0 < err < 2e-5, p


@pytest.mark.parametrize("prec", MIXED)
@pytest.mark.parametrize("order,n,steps", [(3, 16, 40), (1, 16, 40)])
def test_fp32_trajectory_tracks_fp64(solver, order, n, steps, prec):
    """`steps` SSP-RK3 steps of the jittered periodic vortex: mixed-precision state within
    1e-6 relative L2 of the FP64 state; same analytic error; conserved integrals."""
    print(f"[case] {prec} trajectory order={order} n={n} steps={steps}", flush=True)
    mesh = F.Mesh.box(n, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=60)
    U0 = mesh.project(order, "vortex", VORTEX)
    out, cons = {}, {}
    for p in ("fp64", prec):
        solver.set_scheme(order, flux="roe", precision=p).set_mesh(mesh)
        solver.upload(U0)
        dt = solver.stable_dt(0.3)
        c0 = solver.conservation_sums()
        solver.ssprk3(dt, steps)
        cons[p] = (c0, solver.conservation_sums())
        out[p] = (solver.download(), dt)
    (U64, dt), (U32, _) = out["fp64"], out[prec]
    dev = l2_rel(U32, U64)
    print(f"{prec} vs fp64 relative L2 after {steps} steps: {dev:.3e}", flush=True)
    assert dev < 1e-6
    t = steps * dt
    e64 = mesh.analytic_error(order, "vortex", VORTEX, t, U64)
    e32 = mesh.analytic_error(order, "vortex", VORTEX, t, U32)
    assert np.all(np.abs(e32 - e64) <= 1e-3 * np.abs(e64) + 1e-14)
    for p, (c0, c1) in cons.items():
        # relative to the largest conserved integral (the rho w integral of this flow is zero)
        drift = np.abs(np.asarray(c1) - np.asarray(c0)) / np.abs(c0).max()
        assert drift.max() < (1e-13 if p == "fp64" else 1e-11), (p, drift)


@pytest.mark.parametrize("prec", MIXED)
@pytest.mark.parametrize("order,n", [(2, 6), (3, 6), (4, 4)])
def test_mixed_face_paths_agree(oracle, order, n, prec, monkeypatch):
    """Every face path of the mixed stage (default: the face tiles at P2-P4 in FP32; the row-broadcast
    face-flux pass + element kernel; the LDS-staged face-flux pass; the LDS test pass) within 2e-5 of
    the oracle residual, within 3e-6 of the default path's residual (FP32 contractions summed in
    other orders) and within 1e-6 of its two-step state."""
    print(f"[case] mixed face paths {prec} P{order}", flush=True)
    mesh = F.Mesh.box(n, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=5)
    U = mesh.project(order, "vortex", VORTEX)
    U *= 1 + 1e-2 * np.random.default_rng(3).standard_normal(U.shape)
    out = {}
    for path in ("default", "no_tile", "lds", "lds_flux"):
        monkeypatch.delenv("FVDG_FACE_TEST", raising=False)
        monkeypatch.delenv("FVDG_FACE_FLUX", raising=False)
        monkeypatch.delenv("FVDG_FACE_TILE", raising=False)
        if path == "no_tile":
            monkeypatch.setenv("FVDG_FACE_TILE", "off")
        elif path == "lds":
            monkeypatch.setenv("FVDG_FACE_TEST", "lds")
        elif path == "lds_flux":
            monkeypatch.setenv("FVDG_FACE_FLUX", "lds")
        s = F.Solver(0)
        s.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
        s.upload(U)
        R = s.residual()
        s.ssprk3(0.01, 2)
        out[path] = (R, s.download())
        s.close()
    Ro = oracle.residual(mesh.export(), order, U, flux="roe")
    for path, (R, U2) in out.items():
        assert rel_err(R, Ro) < 2e-5, path
        assert rel_err(R, out["default"][0]) < 3e-6, path
        assert rel_err(U2, out["default"][1]) < 1e-6, path


@pytest.fixture(scope="module")
def fp32_faces_solver():
    old = os.environ.get("FVDG_BF16X3_FACES")
    os.environ["FVDG_BF16X3_FACES"] = "0"
    try:
        s = F.Solver(0)
    finally:
        if old is None:
            del os.environ["FVDG_BF16X3_FACES"]
        else:
            os.environ["FVDG_BF16X3_FACES"] = old
    yield s
    s.close()


@pytest.mark.parametrize("order", [2, 3, 4])
def test_bf16x3_face_contractions(fp32_faces_solver, solver, oracle, order):
    """BF16x3's face tiles run the traces and the face test as BF16x3 products (k_face_tile_bf3);
    FVDG_BF16X3_FACES=0 keeps those two in FP32.  Both residuals within 2e-5 of the oracle, not
    equal; 20 SSP-RK3 steps of each within 1e-6 (relative L2) of FP64."""
    print(f"[case] bf16x3 face contractions P{order}", flush=True)
    mesh = F.Mesh.box(4, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=12)
    U = mesh.project(order, "vortex", VORTEX)
    U *= 1 + 1e-2 * np.random.default_rng(order).standard_normal(U.shape)
    Ro = oracle.residual(mesh.export(), order, U, flux="roe")
    R = {}
    for name, s, flux in (("bf3", solver, f"k_face_tile_bf3<{order}>"),
                          ("fp32", fp32_faces_solver, f"k_face_tile<{order},float>")):
        s.set_scheme(order, flux="roe", precision="bf16x3").set_mesh(mesh)
        assert s.stage_kernels()["flux"] == flux
        s.upload(U)
        R[name] = s.residual()
        err = rel_err(R[name], Ro)
        print(f"bf16x3 ({name} faces) residual vs oracle: {err:.3e}", flush=True)
        assert err < 2e-5
    assert not np.array_equal(R["bf3"], R["fp32"])
    U0 = mesh.project(order, "vortex", VORTEX)
    out = {}
    for name, s, prec in (("fp64", solver, "fp64"), ("bf3", solver, "bf16x3"), ("fp32", fp32_faces_solver, "bf16x3")):
        s.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
        s.upload(U0)
        s.ssprk3(0.01, 20)
        out[name] = s.download()
    for name in ("bf3", "fp32"):
        dev = float(np.sqrt(((out[name] - out["fp64"]) ** 2).sum() / (out["fp64"] ** 2).sum()))
        print(f"bf16x3 ({name} faces) vs fp64 after 20 steps: {dev:.2e}", flush=True)
        assert dev < 1e-6
