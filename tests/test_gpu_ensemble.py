"""GPU ensembles: independent problems of uneven size and order stepped together.

Bars: every problem after ensemble steps equals the oracle's SSP-RK3 of that problem alone
(1e-11 relative, FP64); per-problem wave speeds equal a single-problem context's; HIP graph
replays equal eager steps exactly, with host work between the replays; the autotuner leaves
the state as it found it; the FP32-contraction ensemble tracks the FP64 one (1e-6 relative L2).
"""
import numpy as np
import pytest

import fve_dg_amd as F
from conftest import rel_err

pytestmark = pytest.mark.gpu

SPECS = [  # hexes per axis, order, vortex strength, centre (resolved: no blow-up over the steps)
    (3, 1, 2.0, (5.0, 5.0)),
    (4, 3, 3.0, (4.0, 6.0)),
    (3, 2, 2.0, (6.0, 4.5)),
    (4, 4, 2.5, (5.5, 5.0)),
    (3, 1, 1.5, (3.5, 6.5)),
    (3, 3, 2.0, (5.0, 4.0)),
]


def vortex(strength, centre):
    return [1.4, strength, centre[0], centre[1], 1.0, 1.0, 10.0, 10.0]


def build(precision="fp64", flux="roe"):
    meshes = [F.Mesh.box(n, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=40 + i) for i, (n, *_r) in enumerate(SPECS)]
    ens = F.Ensemble(meshes, [s[1] for s in SPECS], flux=flux, precision=precision)
    U0 = [m.project(s[1], "vortex", vortex(s[2], s[3])) for m, s in zip(meshes, SPECS)]
    for p, U in enumerate(U0):
        ens.upload(p, U)
    return meshes, ens, U0


def test_ensemble_matches_oracle_per_problem(oracle):
    print("[case] ensemble vs oracle", flush=True)
    meshes, ens, U0 = build()
    assert sorted(ens.groups()) == [1, 2, 3, 4]
    lam = ens.max_wave_speeds()
    for p, (m, s) in enumerate(zip(meshes, SPECS)):
        solo = F.Solver(0).set_scheme(s[1], flux="roe").set_mesh(m)
        solo.upload(U0[p])
        assert abs(lam[p] - solo.max_wave_speed()) <= 1e-14 * lam[p]
        solo.close()
    h = np.array([m.geometry()["h"].min() * 3.0 for m in meshes])
    dt = np.array([0.3 * h[p] / ((2 * s[1] + 1) * lam[p]) for p, s in enumerate(SPECS)])
    ens.set_dt(dt)
    ens.step(2, graph=False)
    for p, (m, s) in enumerate(zip(meshes, SPECS)):
        ref = oracle.ssprk3(m.export(), s[1], U0[p], dt[p], 2, flux="roe")
        assert rel_err(ens.download(p), ref) < 1e-11, p
    ens.close()


def test_graph_replays_equal_eager_steps():
    print("[case] ensemble graph replay", flush=True)
    _, eager, _ = build()
    _, graph, _ = build()
    dt = np.full(len(SPECS), 0.01)
    for e in (eager, graph):
        e.set_dt(dt)
    eager.step(2, graph=False)
    graph.step(2, graph=True)  # captures one step, replays it twice
    for p in range(len(SPECS)):
        g = graph.download(p)  # host work between replays
        assert np.isfinite(g).all() and np.array_equal(g, eager.download(p)), p
    eager.step(3, graph=False)
    graph.step(3, graph=True)
    for p in range(len(SPECS)):
        assert np.array_equal(graph.download(p), eager.download(p)), p
    # new time steps reach the captured graph through the per-element step array
    dt2 = np.linspace(0.004, 0.009, len(SPECS))
    for e in (eager, graph):
        e.set_dt(dt2)
    eager.step(1, graph=False)
    graph.step(1, graph=True)
    for p in range(len(SPECS)):
        assert np.array_equal(graph.download(p), eager.download(p)), p
    eager.close()
    graph.close()


def test_autotune_keeps_state_and_results(oracle):
    print("[case] ensemble autotune", flush=True)
    meshes, ens, U0 = build()
    ens.set_dt(np.full(len(SPECS), 0.01))
    choice = ens.autotune(trial_steps=1)
    assert set(choice) == {1, 2, 3, 4}
    for p in range(len(SPECS)):
        assert np.array_equal(ens.download(p), U0[p]), p
    ens.step(1, graph=True)
    for p, (m, s) in enumerate(zip(meshes, SPECS)):
        ref = oracle.ssprk3(m.export(), s[1], U0[p], 0.01, 1, flux="roe")
        assert rel_err(ens.download(p), ref) < 1e-11, p
    ens.close()


@pytest.mark.parametrize("tuned", [False, True])
def test_mixed_precision_ensemble_tracks_fp64(tuned):
    """FP32 and BF16x3 contractions in a graph-replayed ensemble (as built, or after the autotuner
    picked each group's kernels: BF16x3 groups at P2-P4 may take their face tiles) stay within 1e-6
    (relative L2) of the FP64 ensemble after 10 steps, problem by problem."""
    print(f"[case] ensemble fp32 / bf16x3 vs fp64, autotuned {tuned}", flush=True)
    out = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        _, ens, _ = build(precision=prec)
        ens.set_dt(np.full(len(SPECS), 0.01))
        if tuned:
            choice = ens.autotune(trial_steps=1)
            print(f"{prec}: {choice}", flush=True)
        ens.step(10, graph=True)
        out[prec] = [ens.download(p) for p in range(len(SPECS))]
        ens.close()
    for prec in ("fp32", "bf16x3"):
        for p in range(len(SPECS)):
            a, b = out[prec][p], out["fp64"][p]
            assert np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()) < 1e-6, (prec, p)


def test_concurrent_groups_equal_serial_groups():
    """order groups on their own streams (forked and joined per step, also inside the graph) give
    the same states bit for bit as the groups one after another on one stream"""
    print("[case] ensemble concurrent vs serial groups", flush=True)
    _, conc, _ = build()
    _, ser, _ = build()
    ser.set_serial(True)
    for e in (conc, ser):
        e.set_dt(np.full(len(SPECS), 0.01))
        e.step(2, graph=True)
        e.step(1, graph=False)
    for p in range(len(SPECS)):
        assert np.array_equal(conc.download(p), ser.download(p)), p
    conc.close()
    ser.close()
