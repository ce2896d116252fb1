"""The long-run validation driver (fve_dg_amd.validate) on small meshes: series recorded at the
chunk ends, the FP32 and BF16x3 contractions tracking FP64 (kinetic energy, enstrophy, conserved
integrals, final state), mass and total energy conserved, the vortex's error against the exact
solution."""
import numpy as np
import pytest

from fve_dg_amd import validate

pytestmark = pytest.mark.gpu


def test_tgv_series():
    print("[case] validate tgv n=4 P2", flush=True)
    r = validate.run("tgv", n=4, order=2, steps=20, every=5, jitter=0.0, precisions=("fp64", "fp32", "bf16x3"))
    for prec in ("fp64", "fp32", "bf16x3"):
        S = np.array(r[prec]["series"])
        assert S.shape == (5, 7) and np.isfinite(S).all()
        assert abs(S[0, 1] - 0.125) < 5e-3  # E_k(0) of the projected field
        # mass and total energy: FP64 to round-off; FP32 contractions sum each side's face-flux
        # differences in FP32, so the two sides of a face cancel to FP32 round-off of those
        tol = 1e-12 if prec == "fp64" else 1e-9
        assert (np.abs(S[:, 3] - S[0, 3]) < tol * S[0, 3]).all(), prec  # mass
        assert (np.abs(S[:, 4] - S[0, 4]) < tol * S[0, 4]).all(), prec  # total energy
        assert S[-1, 1] <= S[0, 1]  # the scheme only dissipates kinetic energy
    for prec in ("fp32", "bf16x3"):
        assert r[f"l2_rel_deviation_{prec}_vs_fp64"] < 1e-7, prec
        assert max(r[f"series_max_rel_deviation_{prec}"].values()) < 1e-7, prec


def test_vortex_errors():
    print("[case] validate vortex n=6 P2", flush=True)
    r = validate.run("vortex", n=6, order=2, steps=10, every=5, jitter=0.1)
    # relative errors of rho, rho u, rho v, rho E (the exact rho w is zero: its relative error is undefined)
    keep = [0, 1, 2, 4]
    e64, e32 = np.array(r["fp64"]["analytic_l2_rel"])[keep], np.array(r["fp32"]["analytic_l2_rel"])[keep]
    assert np.isfinite(e64).all() and e64[0] < 0.05
    assert np.abs(e32 - e64).max() < 1e-6


def test_implicit_steps_track_explicit():
    print("[case] validate implicit n=4 P1", flush=True)
    r = validate.run_implicit(n=4, order=1, steps=2, jitter=0.1, dt_factor=4.0, tol=1e-10)
    assert all(p["picard_iters"] >= 1 and p["linear_iters"] >= 1 for p in r["per_step"])
    ei, ee = np.array(r["analytic_l2_rel_implicit"]), np.array(r["analytic_l2_rel_explicit"])
    assert np.isfinite(ei).all() and ei[0] < 0.2 and r["l2_rel_implicit_vs_explicit"] < 0.05
