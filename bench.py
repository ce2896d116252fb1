#!/usr/bin/env python3
"""Throughput of the explicit DG Euler RK stage on one MI355X.

A step is one SSP-RK3 time step (3 stages; each stage = residual + inverse mass + stage update)
over the whole mesh, with the state resident in HBM.  The metric is DOF-updates per second per
RK stage: 5*N*E*3*K / t over K timed steps (HIP events on the solver stream).

Workloads (--workload):
  vortex_p3_1296k  (default) 60^3 x 6 = 1,296,000 tets, periodic [0,10]^3, P3, Roe, vertices
                   jittered by +-10 % of h: the configuration the FP64-vs-mixed metric is quoted on
  vortex_p3_196k   32^3 x 6 = 196,608 tets, P3, Roe
  vortex_p1_3k     8^3 x 6 = 3,072 tets, P1, Lax-Friedrichs
  tgv_p4_384k      40^3 x 6 = 384,000 tets on [0,2pi]^3, P4, Roe, Taylor-Green vortex (Mach 0.1)
  ensemble_256     256 independent vortex problems (n ~ U{6..24}, P1..P4, strength U(1,5),
                   centre U(3,7)^2, seed 256), HIP-graph replays, autotuned
The vortex is the isentropic vortex (gamma 1.4, strength 5, centre (5,5), free stream (1,1,0),
rho = p = 1 at infinity).  `value` is the FP64 path (the drop-in); `mixed` (FP32 contractions)
and `mixed_bf16x3` time the mixed-precision paths on the same workload from the same initial
state and report their relative L2 deviation from the FP64 state and, for the vortex, every
path's relative L2 error against the exact (translated) vortex.

Roofline (the algorithmic counts of the path, per element and stage, N modes, Q volume and Qf
face points):
  flops  F = 40QN + 145Q + 80QfN + 410Qf + 25N   (volume 40QN + 145Q; faces: traces 40QfN,
         test 40QfN, Riemann 410Qf; update 25N)
  bytes  B = 8 * 5N * (2 + 2/3) + 168            (state read and written, U^n read in two of
         three stages, cell and face geometry)
per kernel: the face-flux kernel F = 40QfN + 410Qf, B = 40N + 88 (its stage input and the face
geometry); the element kernel (volume + face test + update) F = 40QN + 145Q + 40QfN + 25N,
B = 40N (2 + 2/3) + 80.  With face tiles (the mixed path at P2-P4) the face test moves from the
element kernel to the face pass: face pass F = 80QfN + 410Qf, element kernel F = 40QN + 145Q +
25N, bytes as above.  `roofline` is the dominant (longest) kernel; `traffic` is its HBM bytes
per launch measured by rocprofv3 (FETCH_SIZE x 2 + WRITE_SIZE, calibrated on a device copy)
for this workload, read from profiles/hbm_traffic_<workload>.txt (null when absent).  `stage`
gives the same counts for the whole stage, with the measured bytes beside the algorithmic ones.
Peaks: FP64 78.6 TFLOP/s; the mixed paths price their contraction flops at the FP32 matrix peak
(157.3), BF16x3's contractions (the volume and, on its face tiles, the faces) at the bf16 matrix
peak / 6 (six bf16 products per product), and the rest at FP64.

Printed: ONE JSON line.  The CPU baseline (the oracle, a C restatement with the reference's loop
structure, OpenMP) runs on this host: all threads on the full workload for at least one stage,
and one thread on a 12^3 x 6 sample.
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
PEAK_BF16 = 2.5e15     # flop/s (MI355X dense bf16 matrix); BF16x3 spends 6 bf16 products per product

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


def counts(Q, Qf, N):
    """Algorithmic flops / bytes per element and stage: whole stage and per kernel slot."""
    vol_f = 40 * Q * N + 145 * Q
    flux_f = 40 * Qf * N + 410 * Qf
    test_f = 40 * Qf * N
    upd_f = 25 * N
    st = 8 * 5 * N
    return {
        "stage": (vol_f + flux_f + test_f + upd_f, st * (2 + 2 / 3) + 168),
        "volume": (vol_f, st + 80),
        "flux": (flux_f, st + 88),
        "test": (test_f + upd_f, st * (2 + 2 / 3) + 8),
        "elem": (vol_f + test_f + upd_f, st * (2 + 2 / 3) + 80),
        "contraction": 10 * Q * N + 30 * Q * N + 80 * Qf * N,  # flops of the dense contractions
    }


def read_traffic(workload):
    """{kernel-name prefix: HBM bytes per launch} from profiles/hbm_traffic_<workload>.txt (lines
    'name fetch_KiB write_KiB'; bytes = 1024 (2 fetch + write), the calibration of that file)."""
    f = ROOT / "profiles" / f"hbm_traffic_{workload}.txt"
    out = {}
    if f.exists():
        for line in f.read_text().splitlines():
            parts = line.split()
            if len(parts) == 3 and not line.startswith("#"):
                out[parts[0]] = 1024.0 * (2.0 * float(parts[1]) + float(parts[2]))
    return out


def traffic_of(table, kernel_name):
    """measured bytes of a kernel (names compared without spaces), or None"""
    key = kernel_name.replace(" ", "")
    for k, v in table.items():
        if k.replace(" ", "") == key:
            return v
    return None


def roofline(ms, names, E, Q, Qf, N, precision, traffic):
    """Per-kernel and per-stage roofline of one precision's stage (ms from fvdg_time_stages)."""
    c = counts(Q, Qf, N)
    fused = not names.get("volume")
    # face tiles: the face pass also runs the face test, the element kernel the volume and update
    tiles = (names.get("flux") or "").startswith("k_face_tile")
    bf3_faces = (names.get("flux") or "").startswith("k_face_tile_bf3")
    if tiles:
        c["flux"] = (c["flux"][0] + 40 * Qf * N, c["flux"][1])
        c["elem"] = (c["elem"][0] - 40 * Qf * N, c["elem"][1])
    slots = [("volume", ms[1]), ("flux", ms[2]), ("elem" if fused else "test", ms[3])]
    label = {"volume": "volume", "flux": "flux", "test": "test", "elem": "test"}
    kernels = {}
    for slot, t in slots:
        if t <= 0:
            continue
        F, B = c[slot]
        name = names.get(label[slot]) or slot
        if precision == "fp64":
            peak_f = PEAK_FP64
        else:  # contractions at the FP32 matrix peak (BF16x3 volume: bf16 peak / 6), the rest at FP64
            Fc = {"volume": 40 * Q * N, "flux": (80 if tiles else 40) * Qf * N, "test": 40 * Qf * N,
                  "elem": 40 * Q * N + (0 if tiles else 40 * Qf * N)}[slot]
            Fv = 40 * Q * N if (precision == "bf16x3" and slot in ("volume", "elem")) else 0
            if bf3_faces and slot == "flux":
                Fv = Fc  # the face tiles' traces and test as BF16x3 products
            peak_f = F / (Fv / (PEAK_BF16 / 6) + (Fc - Fv) / PEAK_FP32 + (F - Fc) / PEAK_FP64)
        t_s = t * 1e-3
        tf, tb = F * E / peak_f, B * E / PEAK_HBM
        bound = "mfma" if tf >= tb else "hbm"
        measured = traffic_of(traffic, name) if traffic else None
        kernels[name] = {
            "ms": round(t, 4), "bound": bound,
            "flops_per_elem": F, "bytes_per_elem": round(B, 1),
            "achieved_tflops": round(F * E / t_s / 1e12, 3), "peak_tflops": round(peak_f / 1e12, 1),
            "frac_flops": round(F * E / t_s / peak_f, 4),
            "achieved_gbs": round(B * E / t_s / 1e9, 1), "frac_hbm": round(B * E / t_s / PEAK_HBM, 4),
            "measured_bytes": measured,
            "measured_gbs": round(measured / t_s / 1e9, 1) if measured else None,
        }
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    k = kernels[dom]
    Fs, Bs = c["stage"]
    t_st = ms[0] * 1e-3
    measured_stage = sum(v["measured_bytes"] for v in kernels.values()) \
        if all(v["measured_bytes"] for v in kernels.values()) else None
    if precision == "fp64":
        peak_stage = PEAK_FP64
    else:
        Fv = (40 * Q * N if precision == "bf16x3" else 0) + (80 * Qf * N if bf3_faces else 0)
        peak_stage = Fs / (Fv / (PEAK_BF16 / 6) + (c["contraction"] - Fv) / PEAK_FP32 + (Fs - c["contraction"]) / PEAK_FP64)
    mfma = k["bound"] == "mfma"
    return {
        "bound": k["bound"],
        "achieved": k["achieved_tflops"] if mfma else k["achieved_gbs"],
        "peak": k["peak_tflops"] if mfma else PEAK_HBM / 1e9,
        "unit": "TFLOP/s" if mfma else "GB/s",
        "frac": k["frac_flops"] if mfma else k["frac_hbm"],
        "traffic": k["measured_bytes"],
        "kernel": dom, "kernel_ms": k["ms"],
        "counts": "algorithmic flops and bytes per element and stage (see bench.py docstring)",
        "stage": {"ms": round(ms[0], 4), "flops_per_elem": Fs, "bytes_per_elem": round(Bs, 1),
                  "achieved_tflops": round(Fs * E / t_st / 1e12, 3), "peak_tflops": round(peak_stage / 1e12, 1),
                  "frac_flops": round(Fs * E / t_st / peak_stage, 4),
                  "achieved_gbs": round(Bs * E / t_st / 1e9, 1), "frac_hbm": round(Bs * E / t_st / PEAK_HBM, 4),
                  "measured_bytes": measured_stage,
                  "measured_over_algorithmic_bytes": round(measured_stage / (Bs * E), 3) if measured_stage else None},
        "kernels": kernels,
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, U0, order, flux, field, L, seconds_target=20.0):
    """The oracle (C restatement of ExplicitConvection::eval with the reference's loop structure:
    per-point basis evaluation, a 3x3 inverse per (point, test function), OpenMP dynamic schedule
    with atomic scatter) timed on this host: RK stages (residual + inverse mass + update) of the
    full workload with all threads until `seconds_target` (at least one), and of a 12^3 x 6 sample
    with one thread (three stages).  Median stage time each."""
    import numpy as np

    import fve_dg_amd as F
    from oracle.binding import Oracle

    O = Oracle()
    threads = O.max_threads()

    def stages(m, U, th, n_min, budget):
        M = O.mass(m, order)
        times = []
        t_start = time.perf_counter()
        while len(times) < n_min or (time.perf_counter() - t_start < budget and len(times) < 8):
            t0 = time.perf_counter()
            R = O.residual(m, order, U, flux=flux, threads=th)
            U = U + 1e-4 * (-R / M)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), len(times), sum(times)

    N = F.num_basis(order)
    m = mesh.export()
    E = len(m["cells"])
    t_all, n_all, s_all = stages(m, U0, threads, 1, seconds_target)
    small = F.Mesh.box(12, ub=(L, L, L), periodic=7)
    ms_ = small.export()
    Us = small.project(order, field, VORTEX if field == "vortex" else TGV)
    t_one, n_one, s_one = stages(ms_, Us, 1, 3, 0.0)
    Es = len(ms_["cells"])
    return {"value": 5 * N * E / t_all, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{n_all} RK stage(s) of the full {E}-tet workload, P{order} {flux}, oracle/dg_oracle.c "
                      f"(-O3 -mavx2 -mfma -fopenmp) with {threads} OpenMP threads; median stage {t_all:.2f} s of "
                      f"{s_all:.1f} s",
            "single_thread": {"value": 5 * N * Es / t_one, "unit": "DOF-updates/s", "cores": 1,
                              "sample": f"{n_one} RK stages of a 12^3x6 = {Es}-tet periodic box, P{order} {flux}; "
                                        f"median stage {t_one:.2f} s"}}


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
    traffic = read_traffic(args.workload)
    steps_done = args.warmup + 2 * args.steps  # fvdg_time_stages runs two passes of `steps` steps
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
        r = dict(ms=ms, wall=wall, U=U, value=dof / (ms[0] * 1e-3), names=s.stage_kernels())
        if w["field"] == "vortex":
            r["analytic"] = mesh.analytic_error(order, "vortex", params, steps_done * dt, U)
        res[prec] = r
        print(f"[bench] {prec}: {ms[0]:.4f} ms/stage ({ms[1]:.4f} / {ms[2]:.4f} / {ms[3]:.4f}), "
              f"{r['value']:.4g} DOF-updates/s, kernels {r['names']}", file=sys.stderr, flush=True)
    # the same FP64 stages with the limiter (apply_1 + apply_2, point bounds) after each stage
    s.set_scheme(order, flux=flux, precision="fp64")
    s.upload(U0)
    s.set_stage_limiter(F.LIMIT_BOUNDS | F.LIMIT_POSITIVITY)
    ms_lim = s.time_stages(dt, 3 * args.steps)
    s.set_stage_limiter(0)
    s.close()
    f64 = res["fp64"]

    def deviation(r):
        return float(np.sqrt(((r["U"] - f64["U"]) ** 2).sum() / (f64["U"] ** 2).sum()))

    def analytic(r):
        # relative L2 error per variable (rho, rho u, rho v, rho E; the exact rho w is zero)
        return None if "analytic" not in r else [float(r["analytic"][i]) for i in (0, 1, 2, 4)]

    def mixed_block(prec, label):
        r = res[prec]
        return {"precision": label, "value": round(r["value"], 1), "ms_per_step": round(3 * r["ms"][0], 4),
                "speedup_vs_fp64": round(r["value"] / f64["value"], 4),
                "l2_rel_deviation_vs_fp64": deviation(r), "analytic_l2_rel": analytic(r),
                "after_steps": steps_done, "kernels": r["names"],
                "roofline": roofline(r["ms"], r["names"], E, Q, Qf, N, prec, traffic)}

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
                   "stages_per_step": 3, "dt": dt, "cfl": 0.3, "jitter": w["jitter"],
                   "wall_s": round(f64["wall"], 3), "kernels": f64["names"],
                   "analytic_l2_rel": analytic(f64), "after_steps": steps_done},
        "roofline": roofline(f64["ms"], f64["names"], E, Q, Qf, N, "fp64", traffic),
        "mixed": mixed_block("fp32", "FP32 matrix-core contractions (volume interpolation and test, face traces "
                                     "and face test; FP32 accumulation) on differences from each cell's mean, FP32 "
                                     "volume flux perturbation; FP64 Riemann central flux, pressures and positivity "
                                     "(FP32 upwind dissipation), face flux integrals, state, update and sums"),
        "mixed_bf16x3": mixed_block("bf16x3", "volume contractions as 6 bf16 matrix-core products of 3-part splits "
                                              "(FP32 accumulation) of the FP32 path's differences from each "
                                              "cell's mean, face traces and face test likewise (face tiles); "
                                              "FP64 Riemann central flux, pressures, state, update and sums"),
        "limited": {"limiter": "PositiveLimiter apply_1 + apply_2 after every stage (fp64)",
                    "ms_per_step": round(3 * ms_lim[0], 4),
                    "overhead_vs_fp64": round(ms_lim[0] / f64["ms"][0] - 1, 4)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, U0, order, flux, w["field"], L)
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
        else:
            print(f"[bench] ensemble fp32 autotune {tuned}", file=sys.stderr, flush=True)
        if args.warmup > 0:
            ens.step(args.warmup, graph=True)
        ms = ens.time_steps(args.steps, graph=True)
        res[prec] = dict(ms=ms, value=dofs / (ms / 3 * 1e-3), U=[ens.download(p) for p in range(P)], tuned=tuned)
        if prec == "fp64":  # the same graph with the order groups one after another on one stream
            ens.set_serial(True)
            res[prec]["serial_ms"] = ens.time_steps(args.steps, graph=True)
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
                   "concurrent_order_groups": True,
                   "serial_groups_ms_per_step": round(res["fp64"]["serial_ms"], 4),
                   "autotuned": {str(k): v for k, v in res["fp64"]["tuned"].items()},
                   "autotuned_fp32": {str(k): v for k, v in res["fp32"]["tuned"].items()}},
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="vortex_p3_1296k", choices=sorted(WORKLOADS))
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
