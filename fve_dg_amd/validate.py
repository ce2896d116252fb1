"""Long-run validation of the device solver: FP64 against the mixed-precision path, and against
the analytic solution where one exists.

    python -m fve_dg_amd.validate tgv     [--n 40 --order 4 --steps 1000 --every 50]
    python -m fve_dg_amd.validate vortex  [--n 60 --order 3 --steps 500 --every 100 --jitter 0.1]

tgv:    compressible Taylor-Green vortex on the periodic cube [0, 2pi]^3 (gamma 1.4, rho0 = V0 = 1,
        Mach 0.1), Roe flux; every `every` steps the kinetic energy E_k and enstrophy epsilon
        (both per rho0 |Omega|, from fvdg_flow_diagnostics) of each precision; reported: both
        series and their largest relative deviation, and the final states' relative L2 deviation.
vortex: isentropic vortex (strength 5, centre (5, 5), free stream (1, 1, 0)) on the periodic
        cube [0, 10]^3 with jittered vertices, Roe flux; reported: per-variable relative L2 error
        against the translated exact vortex (err_integral of examples/Vortex/VortexBatch3.cpp:
        349-366) for each precision, and the FP32-vs-FP64 relative L2 deviation.
implicit: the vortex advanced by backward-Euler Picard steps (fvdg_implicit_euler_step: GPU assembly
        of the Picard-linearised blocks, block-Jacobi GMRES) with dt = dt_factor x the explicit stable
        step, against SSP-RK3 at the stable step to the same time; reported: per step the Picard and
        GMRES iterations and the wall time, the analytic error of both, and their difference.
The time step is fixed from the initial state (CFL 0.3).  Prints one JSON object on stdout and
a progress line per chunk on stderr.

    python -m fve_dg_amd.validate implicit [--n 16 --order 2 --steps 5 --dt-factor 10 --jitter 0.1]
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import Mesh, Solver, num_basis

VORTEX = [1.4, 5.0, 5.0, 5.0, 1.0, 1.0, 10.0, 10.0]
TGV = [1.4, 1.0, 1.0, 0.1]  # gamma, rho0, V0, Mach
DEFAULTS = {
    "tgv": dict(n=40, order=4, steps=1000, every=50, jitter=0.0),
    "vortex": dict(n=60, order=3, steps=500, every=100, jitter=0.1),
    "implicit": dict(n=16, order=2, steps=5, every=1, jitter=0.1),
}


def run_implicit(n, order, steps, jitter, dt_factor=10.0, tol=1e-8, restart=30, device=0):
    L = 10.0
    mesh = Mesh.box(n, ub=(L, L, L), periodic=7, jitter=jitter, seed=60)
    U0 = mesh.project(order, "vortex", VORTEX)
    s = Solver(device)
    s.set_scheme(order, flux="lf", precision="fp64").set_mesh(mesh)
    s.upload(U0)
    dt_exp = s.stable_dt(0.3)
    dt = dt_factor * dt_exp
    per = []
    for k in range(steps):
        t0 = time.perf_counter()
        st = s.implicit_euler_step(dt, max_picard=100, picard_tol=1e-8, tol=tol, restart=restart, max_iters=5000)
        st["wall_s"] = round(time.perf_counter() - t0, 4)
        per.append(st)
        print(f"[validate] implicit step {k + 1}/{steps}: {st}", file=sys.stderr, flush=True)
    Ui = s.download()
    if not np.isfinite(Ui).all():
        raise RuntimeError("non-finite implicit state")
    t_end = steps * dt
    n_exp = int(round(t_end / dt_exp))
    s.upload(U0)
    t0 = time.perf_counter()
    s.ssprk3(t_end / n_exp, n_exp)
    wall_exp = time.perf_counter() - t0
    Ue = s.download()
    s.close()
    keep = [0, 1, 2, 4]  # the exact rho w is zero: no relative error
    ei = mesh.analytic_error(order, "vortex", VORTEX, t_end, Ui)[keep]
    ee = mesh.analytic_error(order, "vortex", VORTEX, t_end, Ue)[keep]
    return dict(kind="implicit", tets=mesh.n_cells, order=order, dofs=5 * num_basis(order) * mesh.n_cells,
                flux="lf", jitter=jitter, dt=dt, dt_explicit=dt_exp, steps=steps, t_end=t_end,
                gmres_tol=tol, restart=restart, per_step=per,
                implicit_wall_s=round(sum(p["wall_s"] for p in per), 4), explicit_steps=n_exp,
                explicit_wall_s=round(wall_exp, 4), analytic_l2_rel_implicit=ei.tolist(),
                analytic_l2_rel_explicit=ee.tolist(),
                l2_rel_implicit_vs_explicit=float(np.sqrt(((Ui - Ue) ** 2).sum() / (Ue ** 2).sum())))


def run(kind, n, order, steps, every, jitter, flux="roe", cfl=0.3, precisions=("fp64", "fp32"), device=0):
    L = 2 * math.pi if kind == "tgv" else 10.0
    field, params = ("tgv", TGV) if kind == "tgv" else ("vortex", VORTEX)
    mesh = Mesh.box(n, ub=(L, L, L), periodic=7, jitter=jitter, seed=60)
    U0 = mesh.project(order, field, params)
    s = Solver(device)
    s.set_scheme(order, flux=flux, precision="fp64").set_mesh(mesh)
    s.upload(U0)
    dt = s.stable_dt(cfl)
    vol = L ** 3
    out = dict(kind=kind, tets=mesh.n_cells, order=order, dofs=5 * num_basis(order) * mesh.n_cells, flux=flux,
               steps=steps, dt=dt, t_end=steps * dt, jitter=jitter)
    states = {}
    for prec in precisions:
        s.set_scheme(order, flux=flux, precision=prec)
        s.upload(U0)
        series = []

        def record(t):
            d = s.flow_diagnostics()
            series.append([t, d["kinetic_energy"] / (params[1] * vol) if kind == "tgv" else d["kinetic_energy"],
                           d["enstrophy"] / (params[1] * vol) if kind == "tgv" else d["enstrophy"],
                           d["mass"], d["total_energy"], d["min_rho"], d["min_p"]])

        record(0.0)
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(every, steps - done)
            s.ssprk3(dt, k)
            done += k
            record(done * dt)
            print(f"[validate] {kind} {prec}: step {done}/{steps}, E_k {series[-1][1]:.10g}, "
                  f"enstrophy {series[-1][2]:.10g}", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        U = s.download()
        if not np.isfinite(U).all():
            raise RuntimeError(f"non-finite state ({prec})")
        states[prec] = U
        rec = dict(series_columns=["t", "kinetic_energy", "enstrophy", "mass", "total_energy", "min_rho", "min_p"],
                   series=series, wall_s=round(wall, 3))
        if kind == "vortex":
            rec["analytic_l2_rel"] = mesh.analytic_error(order, field, params, steps * dt, U).tolist()
        out[prec] = rec
    s.close()
    for prec in states:
        if prec == "fp64" or "fp64" not in states:
            continue
        a, b = states[prec], states["fp64"]
        out[f"l2_rel_deviation_{prec}_vs_fp64"] = float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
        s64, sm = np.array(out["fp64"]["series"]), np.array(out[prec]["series"])
        out[f"series_max_rel_deviation_{prec}"] = {
            name: float(np.max(np.abs(sm[:, i] - s64[:, i]) / np.maximum(np.abs(s64[:, i]), 1e-300)))
            for i, name in ((1, "kinetic_energy"), (2, "enstrophy"), (3, "mass"), (4, "total_energy"))}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("kind", choices=sorted(DEFAULTS))
    ap.add_argument("--n", type=int)
    ap.add_argument("--order", type=int)
    ap.add_argument("--steps", type=int)
    ap.add_argument("--every", type=int)
    ap.add_argument("--jitter", type=float)
    ap.add_argument("--flux", default="roe", choices=["lf", "roe"])
    ap.add_argument("--dt-factor", type=float, default=10.0, help="implicit: dt / explicit stable dt")
    ap.add_argument("--precisions", default="fp64,fp32", help="tgv/vortex: comma-separated, fp64 first")
    args = ap.parse_args(argv)
    cfg = dict(DEFAULTS[args.kind])
    for k in cfg:
        if getattr(args, k) is not None:
            cfg[k] = getattr(args, k)
    if args.kind == "implicit":
        cfg.pop("every")
        print(json.dumps(run_implicit(dt_factor=args.dt_factor, **cfg)))
    else:
        print(json.dumps(run(args.kind, flux=args.flux, precisions=tuple(args.precisions.split(",")), **cfg)))


if __name__ == "__main__":
    main()
