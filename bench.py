#!/usr/bin/env python3
"""Throughput of the explicit DG Euler RK stage on one MI355X.

A step is one SSP-RK3 time step (3 stages; each stage = residual + inverse mass +
stage update) over the whole mesh, with the state resident in HBM.  The metric is
DOF-updates per second per RK stage: 5*N*E*3*K / t over K timed steps.

Workloads (--workload):
  vortex_p3_196k   (default) 32^3 hexes x 6 = 196,608-tet periodic cube [0,10]^3, P3, Roe
  vortex_p1_3k     8^3 x 6 = 3,072 tets, P1, Lax-Friedrichs
  vortex_p3_1296k  60^3 x 6 = 1,296,000 tets, P3, Roe, vertices jittered by +-10 % of h
  tgv_p4_384k      40^3 x 6 = 384,000 tets on [0,2pi]^3, P4, Roe, Taylor-Green vortex (Mach 0.1)
  ensemble_256     256 independent vortex problems (n ~ U{6..24}, P1..P4, strength U(1,5),
                   centre U(3,7)^2, seed 256), one HIP-graph replay per step, autotuned
The vortex is the isentropic vortex (gamma 1.4, strength 5, centre (5,5), free stream
(1,1,0), rho = p = 1 at infinity).  `value` is the FP64 path (the drop-in); the `mixed`
object times the FP32-contraction path on the same workload and reports its relative L2
deviation from the FP64 state after the same steps.

Printed: ONE JSON line with the metric, the roofline of the dominant kernel (HIP events on
the solver stream), and the CPU oracle timed on a bounded sample on this host's cores.
"""
import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM = 8.0e12      # B/s
PEAK_FP64 = 78.6e12    # flop/s (MI355X dense FP64, vector = matrix)
PEAK_FP32 = 157.3e12   # flop/s (MI355X dense FP32 matrix)

METRIC = "DOF-updates/s per RK stage, P3 tet Euler: FP64 vs mixed; % HBM/MFMA roofline"

WORKLOADS = {
    # name: hexes per axis, order, flux, jitter, field, box edge
    "vortex_p1_3k": dict(n=8, order=1, flux="lf", jitter=0.0, field="vortex", L=10.0),
    "vortex_p3_196k": dict(n=32, order=3, flux="roe", jitter=0.0, field="vortex", L=10.0),
    "vortex_p3_1296k": dict(n=60, order=3, flux="roe", jitter=0.1, field="vortex", L=10.0),
    "tgv_p4_384k": dict(n=40, order=4, flux="roe", jitter=0.0, field="tgv", L=2 * math.pi),
    "ensemble_256": dict(ensemble=256, flux="roe"),
}
VORTEX = [1.4, 5.0, 5.0, 5.0, 1.0, 1.0, 10.0, 10.0]
TGV = [1.4, 1.0, 1.0, 0.1]  # gamma, rho0, V0, Mach

# HBM bytes per launch at the default workload (vortex_p3_196k, FP64) from rocprofv3 FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (factors from a 1 GiB device copy), profiles/dg_stage_p3_196k_fp64.md
TRAFFIC = {
    "vortex_p3_196k": {
        "volume (k_vol_mfma)": (85029 * 2 + 153600) * 1024,
        "face flux (k_face_flux_dpp)": (151599 * 2 + 376357) * 1024,
        "face test + update (k_face_test_mfma)": (623421 * 2 + 176128) * 1024,
    }
}


def test_kernel_name(precision, order):
    """The test pass runs on the matrix cores up to P4 in FP64 and up to P3 in FP32 (unless
    FVDG_FACE_TEST=lds), otherwise on the LDS kernel."""
    mfma = order <= (4 if precision == "fp64" else 3) and os.environ.get("FVDG_FACE_TEST") != "lds"
    return "k_face_test_mfma" if mfma else "k_face_test_lds"


def flux_kernel_name(precision, order):
    """The face-flux pass at P2-P4 holds the coefficients in registers and broadcasts them along
    16-lane rows, with FP64 traces in every precision (unless FVDG_FACE_FLUX=lds); otherwise it
    stages them in LDS."""
    dpp = 2 <= order <= 4 and os.environ.get("FVDG_FACE_FLUX") != "lds"
    return "k_face_flux_dpp" if dpp else "k_face_flux_lds"


def algorithmic(Q, Qf, N, word=8):
    """Algorithmic work per element and stage (FMA = 2 flops).
    volume kernel      40QN + 145Q flops (interpolation 10QN, test 30QN, pointwise flux ~145Q);
                       ~17 flop/B at P3: compute bound, counted against the matrix-core peak
    face-flux kernel   40QfN + 410Qf flops per element (2 faces x [2 traces 10QfN + Riemann ~205Qf])
    face-test kernel   reads U, R, 4 face-flux blocks, U^n (2 of 3 stages), |det|; writes U:
                       8*(5N*(3 + 2/3) + 1) + word*4*Qf*5 bytes, ~3 flop/B: HBM bound."""
    vol_flops = 40 * Q * N + 145 * Q
    flux_flops = 40 * Qf * N + 410 * Qf
    test_bytes = 8 * (5 * N * (3 + 2 / 3) + 1) + word * 4 * Qf * 5
    return vol_flops, flux_flops, test_bytes


def cpu_baseline(workload_n, order, flux, field, L, seconds_target=10.0):
    """The oracle (C restatement, OpenMP, reference loop structure) on a bounded
    sample of the same workload: a periodic box of min(n, 16)^3 x 6 tets at the same
    order and flux; RK stages (residual + inverse mass + update) repeated until
    `seconds_target` of CPU work, median stage time."""
    import numpy as np

    import fve_dg_amd as F
    from oracle.binding import Oracle

    O = Oracle()
    threads = O.max_threads()
    n = min(workload_n, 16)
    mesh = F.Mesh.box(n, ub=(L, L, L), periodic=7)
    m = mesh.export()
    U = mesh.project(order, field, VORTEX if field == "vortex" else TGV)
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


def kernel_roofline(workload, ms, E, Q, Qf, N, precision, order):
    vol_f, flux_f, test_b = algorithmic(Q, Qf, N, 8 if precision == "fp64" else 4)
    peak_mm = PEAK_FP64 if precision == "fp64" else PEAK_FP32

    def rate(work, t_ms):
        return work * E / (t_ms * 1e-3) if t_ms > 0 else 0.0

    kernels = {
        "volume (k_vol_mfma)": {"ms": ms[1], "achieved": rate(vol_f, ms[1]), "peak": peak_mm, "bound": "mfma",
                                "unit": "TFLOP/s", "scale": 1e12},
        f"face flux ({flux_kernel_name(precision, order)})": {
            "ms": ms[2], "achieved": rate(flux_f, ms[2]), "peak": PEAK_FP64, "bound": "mfma", "unit": "TFLOP/s",
            "scale": 1e12},
        f"face test + update ({test_kernel_name(precision, order)})": {
            "ms": ms[3], "achieved": rate(test_b, ms[3]), "peak": PEAK_HBM, "bound": "hbm", "unit": "GB/s",
            "scale": 1e9},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    k = kernels[dom]
    traffic = TRAFFIC.get(workload, {}).get(dom) if precision == "fp64" else None
    return {"bound": k["bound"], "achieved": round(k["achieved"] / k["scale"], 3),
            "peak": round(k["peak"] / k["scale"], 1), "unit": k["unit"],
            "frac": round(k["achieved"] / k["peak"], 4), "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
            "kernel": dom, "kernel_ms": round(k["ms"], 4),
            "all_kernels": {name: {"ms": round(v["ms"], 4), "achieved": round(v["achieved"] / v["scale"], 3),
                                   "unit": v["unit"], "frac": round(v["achieved"] / v["peak"], 4)}
                            for name, v in kernels.items()}}


def run_single(args, w):
    import numpy as np

    import fve_dg_amd as F

    n, order, flux, L = w["n"], w["order"], w["flux"], w["L"]
    params = VORTEX if w["field"] == "vortex" else TGV
    t_setup = time.perf_counter()
    mesh = F.Mesh.box(n, ub=(L, L, L), periodic=7, jitter=w["jitter"], seed=60)
    U0 = mesh.project(order, w["field"], params)
    E = mesh.n_cells
    N = F.num_basis(order)
    _, vw, _, fw = F.quadrature(order)
    Q, Qf = len(vw), len(fw)
    s = F.Solver(0)
    s.set_scheme(order, flux=flux, precision="fp64").set_mesh(mesh)
    s.upload(U0)
    dt = s.stable_dt(0.3)
    print(f"[bench] {args.workload}: E={E} P{order} {flux} dt={dt:.3e} setup {time.perf_counter() - t_setup:.1f}s",
          file=sys.stderr, flush=True)
    dof = 5 * N * E
    res = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        s.set_scheme(order, flux=flux, precision=prec)
        s.upload(U0)
        if args.warmup > 0:
            s.ssprk3(dt, args.warmup)
        t0 = time.perf_counter()
        ms = s.time_stages(dt, 3 * args.steps)  # ms[0]: per-stage average over the timed region
        wall = time.perf_counter() - t0
        U = s.download()
        if not np.isfinite(U).all():
            print(f"[bench] non-finite state after the timed steps ({prec})", file=sys.stderr)
            sys.exit(3)
        res[prec] = dict(ms=ms, wall=wall, U=U, value=dof / (ms[0] * 1e-3))
        print(f"[bench] {prec}: {ms[0]:.4f} ms/stage, {res[prec]['value']:.4g} DOF-updates/s", file=sys.stderr,
              flush=True)
    # the same FP64 stages with the limiter (apply_1 + apply_2, point bounds) after each stage
    s.set_scheme(order, flux=flux, precision="fp64")
    s.upload(U0)
    s.set_stage_limiter(F.LIMIT_BOUNDS | F.LIMIT_POSITIVITY)
    ms_lim = s.time_stages(dt, 3 * args.steps)
    s.set_stage_limiter(0)
    s.close()
    f64, f32, b3 = res["fp64"], res["fp32"], res["bf16x3"]

    def deviation(r):
        return float(np.sqrt(((r["U"] - f64["U"]) ** 2).sum() / (f64["U"] ** 2).sum()))

    dev = deviation(f32)
    out = {
        "metric": METRIC,
        "value": round(f64["value"], 1),
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(3 * f64["ms"][0], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic ({'isentropic vortex' if w['field'] == 'vortex' else 'Taylor-Green vortex'} projected "
                f"on the mesh)",
        "config": {"workload": args.workload, "tets": E, "order": order, "modes": N, "flux": flux, "dofs": dof,
                   "stages_per_step": 3, "dt": dt, "cfl": 0.3, "wall_s": round(f64["wall"], 3)},
        "roofline": kernel_roofline(args.workload, f64["ms"], E, Q, Qf, N, "fp64", order),
        "mixed": {"precision": "fp32 contractions, fp64 fluxes/update/state",
                  "value": round(f32["value"], 1), "ms_per_step": round(3 * f32["ms"][0], 4),
                  "speedup_vs_fp64": round(f32["value"] / f64["value"], 4),
                  "l2_rel_deviation_vs_fp64": dev, "after_steps": args.warmup + args.steps,
                  "roofline": kernel_roofline(args.workload, f32["ms"], E, Q, Qf, N, "fp32", order)},
        "mixed_bf16x3": {"precision": "volume contractions as 6 bf16 MFMA products of 3-part splits (fp32 "
                                      "accumulation), face contractions fp32, fp64 fluxes/update/state",
                         "value": round(b3["value"], 1), "ms_per_step": round(3 * b3["ms"][0], 4),
                         "speedup_vs_fp64": round(b3["value"] / f64["value"], 4),
                         "l2_rel_deviation_vs_fp64": deviation(b3), "after_steps": args.warmup + args.steps,
                         "volume_kernel_ms": round(b3["ms"][1], 4)},
        "limited": {"limiter": "PositiveLimiter apply_1 + apply_2 after every stage (fp64)",
                    "ms_per_step": round(3 * ms_lim[0], 4),
                    "overhead_vs_fp64": round(ms_lim[0] / f64["ms"][0] - 1, 4)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, order, flux, w["field"], L)
    return out


def run_ensemble(args, w):
    import numpy as np

    import fve_dg_amd as F

    rng = np.random.default_rng(256)
    P = w["ensemble"]
    ns = rng.integers(6, 25, P)
    orders = rng.integers(1, 5, P)
    strength = rng.uniform(1, 5, P)
    centre = rng.uniform(3, 7, (P, 3))
    t_setup = time.perf_counter()

    def make(p):
        m = F.Mesh.box(int(ns[p]), ub=(10, 10, 10), periodic=7)
        U = m.project(int(orders[p]), "vortex", [1.4, strength[p], centre[p, 0], centre[p, 1], 1.0, 1.0, 10.0, 10.0])
        return m, U

    with ThreadPoolExecutor(16) as ex:
        built = list(ex.map(make, range(P)))
    meshes = [b[0] for b in built]
    dofs = sum(5 * F.num_basis(int(o)) * m.n_cells for m, o in zip(meshes, orders))
    tets = sum(m.n_cells for m in meshes)
    res = {}
    for prec in ("fp64", "fp32"):
        ens = F.Ensemble(meshes, orders, flux=w["flux"], precision=prec)
        for p in range(P):
            ens.upload(p, built[p][1])
        lam = ens.max_wave_speeds()
        h = np.array([3.0 * m.geometry()["h"].min() for m in meshes]) if prec == "fp64" else h
        dt = 0.3 * h / ((2 * orders + 1) * lam)
        ens.set_dt(dt)
        tuned = ens.autotune(trial_steps=2)
        if prec == "fp64":
            print(f"[bench] ensemble: {P} problems, {tets} tets, {dofs} DOF, setup "
                  f"{time.perf_counter() - t_setup:.1f}s, autotune {tuned}", file=sys.stderr, flush=True)
        if args.warmup > 0:
            ens.step(args.warmup, graph=True)
        ms = ens.time_steps(args.steps, graph=True)
        res[prec] = dict(ms=ms, value=dofs / (ms / 3 * 1e-3), U=[ens.download(p) for p in range(P)], tuned=tuned)
        ens.close()
        print(f"[bench] {prec}: {ms:.3f} ms/step, {res[prec]['value']:.4g} DOF-updates/s", file=sys.stderr,
              flush=True)
    num = sum(((a - b) ** 2).sum() for a, b in zip(res["fp32"]["U"], res["fp64"]["U"]))
    den = sum((b ** 2).sum() for b in res["fp64"]["U"])
    return {
        "metric": METRIC,
        "value": round(res["fp64"]["value"], 1),
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["fp64"]["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (256 isentropic vortices projected on their meshes)",
        "config": {"workload": args.workload, "problems": P, "tets": tets, "dofs": dofs,
                   "orders": {int(o): int((orders == o).sum()) for o in sorted(set(orders.tolist()))},
                   "hexes_per_axis": [int(ns.min()), int(ns.max())], "stages_per_step": 3, "hip_graph": True,
                   "autotuned": {str(k): v for k, v in res["fp64"]["tuned"].items()}},
        "roofline": None,
        "mixed": {"precision": "fp32 contractions, fp64 fluxes/update/state", "value": round(res["fp32"]["value"], 1),
                  "ms_per_step": round(res["fp32"]["ms"], 4),
                  "speedup_vs_fp64": round(res["fp32"]["value"] / res["fp64"]["value"], 4),
                  "l2_rel_deviation_vs_fp64": float(math.sqrt(num / den)),
                  "after_steps": args.warmup + args.steps + 1},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="vortex_p3_196k", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="before the run, copy MIB MiB device-to-device once (a plain copy kernel of known "
                         "size, to calibrate the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU path is a design (element partition + halo "
              "exchange), not built; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy > 0:
        import torch

        src = torch.ones(args.calibrate_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    w = WORKLOADS[args.workload]
    out = run_ensemble(args, w) if "ensemble" in w else run_single(args, w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
