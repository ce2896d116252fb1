#!/usr/bin/env python3
"""Throughput of the explicit DG Euler RK stage on one MI355X.

A step is one SSP-RK3 time step (3 stages; each stage = residual + inverse mass +
stage update) over the whole mesh, with the state resident in HBM.  The metric
is DOF-updates per second per RK stage: 5*N*E*3*K / t over K timed steps.

Default workload: the 196,608-tet periodic cube (32^3 hexes x 6 tets, [0,10]^3)
at P3 (N = 20 modes x 5 variables) with the Roe flux in FP64, isentropic vortex
(gamma 1.4, strength 5, centre (5,5), free stream (1,1,0)) as the state.

Printed: ONE JSON line with the metric, the roofline of the dominant kernel
(measured with HIP events on the solver stream), and the CPU oracle timed on
a bounded sample on this host's cores.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM = 8.0e12      # B/s
PEAK_FP64 = 78.6e12    # flop/s (MI355X dense FP64, vector = matrix)

WORKLOADS = {
    # name: (hexes per axis, order, flux, jitter)
    "vortex_p3_196k": (32, 3, "roe", 0.0),
    "vortex_p3_1296k": (60, 3, "roe", 0.1),
    "vortex_p1_3k": (8, 1, "lf", 0.0),
}
VORTEX = [1.4, 5.0, 5.0, 5.0, 1.0, 1.0, 10.0, 10.0]

# HBM bytes per launch at the default workload (vortex_p3_196k) from rocprofv3 FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (factors from a 1 GiB device copy), profiles/dg_stage_p3_196k_fp64.md
TRAFFIC = {
    "vortex_p3_196k": {
        "volume (k_vol_mfma)": (85037 * 2 + 153600) * 1024,
        "face flux (k_face_flux_lds)": (167752 * 2 + 368656) * 1024,
        "face test + update (k_face_test_lds)": (476464 * 2 + 164627) * 1024,
    }
}


def algorithmic(order, Q, Qf, N):
    """Algorithmic work per element and stage (FMA = 2 flops, 8-byte words).
    volume kernel      40QN + 145Q flops (interpolation 10QN, test 30QN, pointwise flux ~145Q);
                       intensity ~17 flop/B at P3: FP64-compute bound
    face-flux kernel   40QfN + 410Qf flops (2 faces x [2 traces 10QfN + Riemann ~205Qf]);
                       ~11 flop/B: at the FP64 ridge, counted against FP64
    face-test kernel   reads U, R, 4 face-flux blocks, U^n (2 of 3 stages), |det|; writes U:
                       8*(5N*(3 + 2/3) + 4*Qf*5 + 1) bytes, ~3 flop/B: HBM bound."""
    vol_flops = 40 * Q * N + 145 * Q
    flux_flops = 40 * Qf * N + 410 * Qf
    test_bytes = 8 * (5 * N * (3 + 2 / 3) + 4 * Qf * 5 + 1)
    return vol_flops, flux_flops, test_bytes


def cpu_baseline(workload_n, order, flux, seconds_target=10.0):
    """The oracle (C restatement, OpenMP, reference loop structure) on a bounded
    sample of the same workload: a 16^3 x 6 = 24,576-tet periodic box at the
    same order and flux; RK stages (residual + inverse mass + update) repeated
    until `seconds_target` of CPU work, median stage time."""
    import numpy as np

    import fve_dg_amd as F
    from oracle.binding import Oracle

    O = Oracle()
    threads = O.max_threads()
    n = 16 if workload_n >= 16 else workload_n
    mesh = F.Mesh.box(n, ub=(10, 10, 10), periodic=7)
    m = mesh.export()
    U = mesh.project(order, "vortex", VORTEX)
    M = O.mass(m, order)
    E = len(m["cells"])
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or time.perf_counter() - t_start < seconds_target:
        t0 = time.perf_counter()
        R = O.residual(m, order, U, flux=flux, threads=threads)
        U = U + 1e-4 * (-R / M)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    dof = 5 * F.num_basis(order) * E
    return {"value": dof / t, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} RK stages on a {n}^3x6 = {E}-tet periodic box, P{order} {flux}, "
                      f"oracle/dg_oracle.c with {threads} OpenMP threads; median stage {t:.3f} s of "
                      f"{sum(times):.1f} s total"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="vortex_p3_196k", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="contraction precision: fp64 (the drop-in) or fp32 (mixed: FP32 volume/face "
                         "contractions, FP64 fluxes, update and state)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="before the run, copy MIB MiB device-to-device once (a plain copy kernel of known "
                         "size, to calibrate the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU path is a design (element partition + halo "
              "exchange), not built; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import fve_dg_amd as F

    if args.calibrate_copy > 0:
        import torch

        src = torch.ones(args.calibrate_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    n, order, flux, jitter = WORKLOADS[args.workload]
    t_setup = time.perf_counter()
    mesh = F.Mesh.box(n, ub=(10, 10, 10), periodic=7, jitter=jitter, seed=60)
    U = mesh.project(order, "vortex", VORTEX)
    E = mesh.n_cells
    N = F.num_basis(order)
    vp, vw, fp, fw = F.quadrature(order)
    Q, Qf = len(vw), len(fw)
    s = F.Solver(0)
    s.set_scheme(order, flux=flux, precision=args.precision).set_mesh(mesh)
    s.upload(U)
    h = mesh.geometry()["h"].min() * 3.0
    dt = 0.3 * h / ((2 * order + 1) * s.max_wave_speed())
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] {args.workload}: E={E} P{order} {flux} dt={dt:.3e} setup {setup_s:.1f}s", file=sys.stderr,
          flush=True)

    if args.warmup > 0:
        s.ssprk3(dt, args.warmup)
    t0 = time.perf_counter()
    ms = s.time_stages(dt, 3 * args.steps)   # ms[0]: per-stage average over the timed region
    wall = time.perf_counter() - t0
    stage_ms = ms[0]
    dof = 5 * N * E
    value = dof / (stage_ms * 1e-3)

    vol_f, flux_f, test_b = algorithmic(order, Q, Qf, N)

    def rate(work, t_ms):
        return work * E / (t_ms * 1e-3) if t_ms > 0 else 0.0

    kernels = {
        "volume (k_vol_mfma)": {"ms": ms[1], "achieved": rate(vol_f, ms[1]), "peak": PEAK_FP64, "bound": "mfma",
                                "unit": "TFLOP/s", "scale": 1e12},
        "face flux (k_face_flux_lds)": {"ms": ms[2], "achieved": rate(flux_f, ms[2]), "peak": PEAK_FP64,
                                        "bound": "mfma", "unit": "TFLOP/s", "scale": 1e12},
        "face test + update (k_face_test_lds)": {"ms": ms[3], "achieved": rate(test_b, ms[3]), "peak": PEAK_HBM,
                                                 "bound": "hbm", "unit": "GB/s", "scale": 1e9},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    k = kernels[dom]
    roofline = {"bound": k["bound"], "achieved": round(k["achieved"] / k["scale"], 3),
                "peak": round(k["peak"] / k["scale"], 1), "unit": k["unit"],
                "frac": round(k["achieved"] / k["peak"], 4), "traffic": TRAFFIC.get(args.workload, {}).get(dom),
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
                "kernel": dom, "kernel_ms": round(k["ms"], 4),
                "all_kernels": {name: {"ms": round(v["ms"], 4), "achieved": round(v["achieved"] / v["scale"], 3),
                                       "unit": v["unit"], "frac": round(v["achieved"] / v["peak"], 4)}
                                for name, v in kernels.items()}}
    out = {
        "metric": "DOF-updates/s per RK stage, P3 tet Euler (FP64)",
        "value": round(value, 1),
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(3 * stage_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (isentropic vortex projected on the mesh)",
        "config": {"workload": args.workload, "tets": E, "order": order, "modes": N, "flux": flux,
                   "dofs": dof, "stages_per_step": 3, "precision": "fp64", "wall_s": round(wall, 3)},
        "roofline": roofline,
    }
    Ug = s.download()
    if not np.isfinite(Ug).all():
        print("[bench] non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, order, flux)
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
