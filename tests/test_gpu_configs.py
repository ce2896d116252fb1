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
* C3 at full size: the 1,296,000-tet jittered cube's FP64 (1e-12) and FP32-contraction (2e-5) residuals
  against the oracle's (trajectories at 16^3 x 6 in test_gpu_mixed.py, 23 steps at full size in the
  benchmark's deviation figure);
* C4's draw: the first 12 of the 256 problems (seed 256) as one graph-replayed, autotuned ensemble,
  one step, each problem against the oracle (1e-11)."""
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


def test_c5_protocol_full_size():
    """C5's protocol: the inviscid Taylor-Green vortex on 40^3 x 6 = 384,000 tets at P4 (Roe, CFL 0.3),
    1,000 SSP-RK3 steps in FP64 and BF16x3 (FP32 as well): the integrated kinetic-energy and
    enstrophy series every 100 steps within 1e-8 / 1e-7 (relative) of the FP64 series, mass and
    total energy within 1e-12, the final state within 1e-6 (relative L2)."""
    print("[case] C5: TGV 40^3x6 P4 roe, 1000 steps, fp64/fp32/bf16x3 series", flush=True)
    r = validate.run("tgv", n=40, order=4, steps=1000, every=100, jitter=0.0, precisions=("fp64", "fp32", "bf16x3"))
    assert r["tets"] == 384000
    S64 = np.array(r["fp64"]["series"])
    assert S64.shape == (11, 7) and np.isfinite(S64).all()
    # mass and total energy conserved; at t ~ 0.1 the kinetic energy has barely changed (it trades
    # with internal energy in compressible flow: no sign is fixed)
    assert abs(S64[-1, 3] / S64[0, 3] - 1) < 1e-12 and abs(S64[-1, 4] / S64[0, 4] - 1) < 1e-12
    assert abs(S64[-1, 1] / S64[0, 1] - 1) < 1e-5
    for prec in ("fp32", "bf16x3"):
        dev = r[f"series_max_rel_deviation_{prec}"]
        print(f"C5 {prec}: series deviation {dev}, state {r[f'l2_rel_deviation_{prec}_vs_fp64']:.2e}", flush=True)
        assert dev["kinetic_energy"] < 1e-8 and dev["enstrophy"] < 1e-7, prec
        # both keep the cell means' face sums in FP64 (face tiles)
        assert dev["mass"] < 1e-12 and dev["total_energy"] < 1e-12, prec
        assert r[f"l2_rel_deviation_{prec}_vs_fp64"] < 1e-6, prec


def test_c3_residual_full_size(solver, oracle):
    """C3 at full size: the 60^3 x 6 = 1,296,000-tet jittered periodic cube, P3 Roe; the FP64 residual
    within 1e-12 and the FP32-contraction residual within 2e-5 of the oracle's (all host threads)."""
    print("[case] C3: 60^3x6 jittered P3 roe residual against the oracle", flush=True)
    mesh = F.Mesh.box(60, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=60)
    assert mesh.n_cells == 1296000
    U = mesh.project(3, "vortex", VORTEX)
    Ro = oracle.residual(mesh.export(), 3, U, flux="roe", threads=oracle.max_threads())
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-5)):
        solver.set_scheme(3, flux="roe", precision=prec).set_mesh(mesh)
        solver.upload(U)
        err = rel_err(solver.residual(), Ro)
        print(f"C3 {prec} residual vs oracle: {err:.2e}", flush=True)
        assert err < tol


def test_c3_protocol_500_steps(solver):
    """C3's protocol: the 1,296,000-tet jittered cube, P3 Roe, 500 SSP-RK3 steps at CFL 0.3 in FP64
    and with the FP32 contractions.  The mixed state stays within 1e-6 (relative L2) of the FP64
    state, its conserved integrals track FP64's, and both runs carry the same error against the
    translated analytic vortex (per variable, relative L2 at the volume points)."""
    print("[case] C3: 60^3x6 jittered P3 roe, 500 steps fp64 and fp32", flush=True)
    mesh = F.Mesh.box(60, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=60)
    U0 = mesh.project(3, "vortex", VORTEX)
    solver.set_scheme(3, flux="roe", precision="fp64").set_mesh(mesh)
    solver.upload(U0)
    dt = solver.stable_dt(0.3)
    steps = 500
    out, err, sums = {}, {}, {}
    for prec in ("fp64", "fp32"):
        solver.set_scheme(3, flux="roe", precision=prec)
        solver.upload(U0)
        s0 = np.asarray(solver.conservation_sums())
        for _ in range(5):
            solver.ssprk3(dt, steps // 5)
            print(f"C3 {prec}: {_ + 1}/5 of {steps} steps", flush=True)
        out[prec] = solver.download()
        sums[prec] = (np.asarray(solver.conservation_sums()) - s0) / np.abs(s0).max()
        err[prec] = mesh.analytic_error(3, "vortex", VORTEX, steps * dt, out[prec])
        print(f"C3 {prec}: analytic error {err[prec]}, drift {sums[prec]}", flush=True)
    assert np.isfinite(out["fp32"]).all()
    dev = float(np.sqrt(((out["fp32"] - out["fp64"]) ** 2).sum() / (out["fp64"] ** 2).sum()))
    print(f"C3 500 steps: fp32 vs fp64 relative L2 {dev:.2e}", flush=True)
    assert dev < 1e-6
    assert np.abs(sums["fp64"]).max() < 1e-12 and np.abs(sums["fp32"]).max() < 1e-11
    e64, e32 = np.asarray(err["fp64"])[[0, 1, 2, 4]], np.asarray(err["fp32"])[[0, 1, 2, 4]]
    assert (e64 < 1e-4).all()  # the P3 scheme resolves the vortex on this mesh
    assert np.allclose(e32, e64, rtol=1e-3, atol=0)


def test_c4_protocol_full_size():
    """C4's protocol: the 256-problem draw (seed 256: n ~ U{6..24} hexes per axis, P1..P4, strength
    U(1,5), centre U(3,7)), one ensemble per precision (order groups, HIP graph, autotuned), 100
    SSP-RK3 steps each at its own stable dt: the FP32-contraction ensemble within 1e-6 (relative L2)
    of the FP64 ensemble for every problem, and both with the same analytic vortex error."""
    from concurrent.futures import ThreadPoolExecutor

    print("[case] C4: 256 problems, 100 steps, fp64 and fp32 ensembles", flush=True)
    rng = np.random.default_rng(256)
    P = 256
    ns, orders = rng.integers(6, 25, P), rng.integers(1, 5, P)
    strength, centre = rng.uniform(1, 5, P), rng.uniform(3, 7, (P, 3))
    params = [[1.4, strength[p], centre[p, 0], centre[p, 1], 1.0, 1.0, 10.0, 10.0] for p in range(P)]

    def make(p):
        m = F.Mesh.box(int(ns[p]), ub=(10, 10, 10), periodic=7)
        return m, m.project(int(orders[p]), "vortex", params[p])

    with ThreadPoolExecutor(16) as ex:
        built = list(ex.map(make, range(P)))
    meshes = [b[0] for b in built]
    h = np.array([3.0 * m.geometry()["h"].min() for m in meshes])
    steps, out, dt = 100, {}, None
    for prec in ("fp64", "fp32"):
        ens = F.Ensemble(meshes, [int(o) for o in orders], flux="roe", precision=prec)
        try:
            for p in range(P):
                ens.upload(p, built[p][1])
            if dt is None:
                dt = 0.3 * h / ((2 * orders + 1) * ens.max_wave_speeds())
            ens.set_dt(dt)
            ens.autotune(trial_steps=1)
            for p in range(P):  # the autotuner's trials restore the state; start from U0 regardless
                ens.upload(p, built[p][1])
            ens.step(steps, graph=True)
            out[prec] = [ens.download(p) for p in range(P)]
        finally:
            ens.close()
        print(f"C4 {prec}: {steps} steps done", flush=True)
    dev = np.array([np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()) for a, b in zip(out["fp32"], out["fp64"])])
    print(f"C4: fp32 vs fp64 per-problem relative L2, max {dev.max():.2e}, median {np.median(dev):.2e}", flush=True)
    assert all(np.isfinite(u).all() for u in out["fp32"])
    assert dev.max() < 1e-6
    for p in range(0, P, 32):
        e64 = meshes[p].analytic_error(int(orders[p]), "vortex", params[p], steps * dt[p], out["fp64"][p])[[0, 1, 2, 4]]
        e32 = meshes[p].analytic_error(int(orders[p]), "vortex", params[p], steps * dt[p], out["fp32"][p])[[0, 1, 2, 4]]
        assert np.allclose(e32, e64, rtol=1e-3, atol=0), (p, e32, e64)


def test_c4_subset_matches_oracle(oracle):
    """C4's draw (seed 256: n ~ U{6..24}, P1..P4, strength U(1,5), centre U(3,7)): the first 12
    problems as one ensemble (graph-replayed, autotuned), one SSP-RK3 step, each problem against the
    oracle's step of that problem alone (1e-11)."""
    print("[case] C4: first 12 problems of the 256-problem draw against the oracle", flush=True)
    rng = np.random.default_rng(256)
    P = 256
    ns, orders = rng.integers(6, 25, P), rng.integers(1, 5, P)
    strength, centre = rng.uniform(1, 5, P), rng.uniform(3, 7, (P, 3))
    K = 12
    meshes = [F.Mesh.box(int(ns[p]), ub=(10, 10, 10), periodic=7) for p in range(K)]
    params = [[1.4, strength[p], centre[p, 0], centre[p, 1], 1.0, 1.0, 10.0, 10.0] for p in range(K)]
    U0 = [meshes[p].project(int(orders[p]), "vortex", params[p]) for p in range(K)]
    ens = F.Ensemble(meshes, [int(o) for o in orders[:K]], flux="roe")
    try:
        for p in range(K):
            ens.upload(p, U0[p])
        lam = ens.max_wave_speeds()
        h = np.array([3.0 * m.geometry()["h"].min() for m in meshes])
        dt = 0.3 * h / ((2 * orders[:K] + 1) * lam)
        ens.set_dt(dt)
        ens.autotune(trial_steps=1)
        ens.step(1, graph=True)
        th = oracle.max_threads()
        for p in range(K):
            ref = oracle.ssprk3(meshes[p].export(), int(orders[p]), U0[p], dt[p], 1, flux="roe", threads=th)
            err = rel_err(ens.download(p), ref)
            assert err < 1e-11, (p, int(ns[p]), int(orders[p]), err)
    finally:
        ens.close()
