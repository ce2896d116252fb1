"""The C++ drop-in headers (include/DG/DG_Schemes/ExplicitConvection.h, ExplicitSSPRK3.h)
used the way the reference's drivers use ExplicitConvection: tests/cpp/explicit_convection_check
runs on the GPU and writes its inputs and results; they are checked here against the oracle.

* residual and each part (cells / internals / boundarys) within 1e-12 of the oracle, with
  Dirichlet data from the application hooks at t = 0.3 and Pseudo3DZ faces;
* eval = eval_cells + eval_internals + eval_boundarys;
* stable_dt = 0.3 h_min / ((2p+1) lambda_max) recomputed from the traces;
* two driver steps with Dirichlet data refreshed at the stage times t, t+dt, t+dt/2 equal the
  Shu-Osher composition of oracle residuals at those times;
* a periodic jittered mesh with gamma = 5/3 (MonatomicFluxC) and the Roe flux;
* PositiveLimiter<2> (constructMinMax + apply) and PositiveLimiter<2, QuadC, true>
  (constructMinMax + apply_1 + apply_2) on a sharpened vortex equal the oracle's limiter;
* one backward-Euler system built as the reference's implicit drivers build it
  (ImplicitConvection<1>::assemble + mass/dt blocks + finalize + EigenSparseSolver::DGMRES) equals
  the dense solve of the oracle's assembly, and BlockSparseMatrix::multiply equals its product.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import rel_err

EXE = Path(__file__).resolve().parent / "cpp" / "explicit_convection_check"
VORTEX_REF = [1.4, 5.0, 4.0, 4.5, 1.0, 0.5, 0.0, 0.0]  # examples/Vortex/problem.h:7-14


class Reader:
    def __init__(self, path):
        self.b = Path(path).read_bytes()
        self.o = 0

    def i64(self, n=None):
        k = 1 if n is None else n
        a = np.frombuffer(self.b, np.int64, k, self.o)
        self.o += 8 * k
        return int(a[0]) if n is None else a.copy()

    def f64(self, n):
        a = np.frombuffer(self.b, np.float64, n, self.o)
        self.o += 8 * n
        return a.copy()

    def mesh(self):
        npnt, nc, nf = self.i64(), self.i64(), self.i64()
        pts = self.f64(3 * npnt).reshape(npnt, 3)
        cells = self.i64(4 * nc).reshape(nc, 4)
        fr = self.i64(8 * nf).reshape(nf, 8)
        bc = self.i64(nf).astype(np.uint8)
        return dict(points=pts, cells=cells, face_nodes=fr[:, 0:3], face_nodes_r=fr[:, 3:6], face_cells=fr[:, 6:8]), bc


def face_traces(oracle, m, og, order, U):
    """traces of U at every face point, both sides (side 1 = NaN on boundary faces)"""
    _, _, fp, _ = F.quadrature(order)
    N = F.num_basis(order)
    nf = len(m["face_nodes"])
    out = np.full((nf, 2, len(fp), 5), np.nan)
    for f in range(nf):
        for side in (0, 1):
            c = m["face_cells"][f, side]
            if c < 0:
                continue
            vt = og["vert"][f, side]
            C = U[c].reshape(N, 5)
            for g, (u, v) in enumerate(fp):
                xi = np.zeros(4)
                xi[vt[0]] += 1 - u - v
                xi[vt[1]] += u
                xi[vt[2]] += v
                phi, _ = oracle.basis(order, *xi[1:])
                out[f, side, g] = phi @ C
    return out


def dirichlet_states(oracle, m, og, order, bc, t):
    _, _, fp, _ = F.quadrature(order)
    rows = []
    for f in np.nonzero((m["face_cells"][:, 1] < 0) & (bc == 0))[0]:
        X = m["points"][m["cells"][m["face_cells"][f, 0]]]
        vt = og["vert"][f, 0]
        for u, v in fp:
            lam = np.zeros(4)
            lam[vt[0]] += 1 - u - v
            lam[vt[1]] += u
            lam[vt[2]] += v
            rows.append(oracle.vortex_state(VORTEX_REF, t, lam @ X))
    return np.array(rows)


def test_check_program_is_built():
    assert EXE.exists(), "tests/cpp/explicit_convection_check is built by __graft_entry__.build()"


@pytest.mark.skipif(F.gpu_available(), reason="checks the no-device error path")
def test_check_program_fails_loudly_without_a_device(tmp_path):
    r = subprocess.run([str(EXE), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "fvdg_create" in r.stderr


@pytest.mark.gpu
def test_cpp_dropin_matches_oracle(oracle, tmp_path):
    out = tmp_path / "cpp_check.bin"
    print("[case] tests/cpp/explicit_convection_check", flush=True)
    r = subprocess.run([str(EXE), str(out)], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    rd = Reader(out)
    assert rd.i64() == 5
    # ---- Dirichlet / Pseudo3DZ boxes, P1 LF and P3 Roe
    for _ in range(2):
        order, flux_id = rd.i64(), rd.i64()
        flux = "roe" if flux_id == 1 else "lf"
        m, bc = rd.mesh()
        E, N = len(m["cells"]), F.num_basis(order)
        U, R, Rc, Ri, Rb = (rd.f64(E * 5 * N).reshape(E, -1) for _ in range(5))
        dt = rd.f64(1)[0]
        U2, M = rd.f64(E * 5 * N).reshape(E, -1), rd.f64(E * 5 * N).reshape(E, -1)
        og = oracle.geometry(m)
        bnd = m["face_cells"][:, 1] < 0
        assert (bc[bnd] == 5).any() and (bc[bnd] == 0).any()
        t = 0.3
        UD = lambda tt: dirichlet_states(oracle, m, og, order, bc, tt)
        args = dict(flux=flux, bc=bc)
        assert rel_err(R, oracle.residual(m, order, U, UD=UD(t), **args)) < 1e-12
        for parts, Rp in ((1, Rc), (2, Ri), (4, Rb)):
            assert rel_err(Rp, oracle.residual(m, order, U, UD=UD(t), parts=parts, **args)) < 1e-12, parts
        assert rel_err(Rc + Ri + Rb, R) < 1e-14
        # mass and the time step rule
        Mo = oracle.mass(m, order)
        assert rel_err(M, Mo) < 1e-14
        area_sum = np.zeros(E)
        for f, (c0, c1) in enumerate(m["face_cells"]):
            area_sum[c0] += og["area"][f]
            if c1 >= 0:
                area_sum[c1] += og["area"][f]
        h_min = (3 * (og["det"] / 6) / area_sum).min()
        tr = face_traces(oracle, m, og, order, U)
        lam = max(oracle.wave_speed(u, u) for u in tr.reshape(-1, 5) if np.isfinite(u).all())
        assert abs(dt - 0.3 * h_min / ((2 * order + 1) * lam)) < 1e-13 * dt
        # two steps, Dirichlet data at the stage times
        L = lambda V, tt: -oracle.residual(m, order, V, UD=UD(tt), **args) / Mo
        V = U.copy()
        for k in range(2):
            t0 = t + k * dt
            V1 = V + dt * L(V, t0)
            V2 = 0.75 * V + 0.25 * (V1 + dt * L(V1, t0 + dt))
            V = V / 3 + 2 / 3 * (V2 + dt * L(V2, t0 + 0.5 * dt))
        assert rel_err(U2, V) < 1e-11
    # ---- periodic jittered cube, gamma 5/3, P2 Roe
    order, flux_id = rd.i64(), rd.i64()
    m, bc = rd.mesh()
    E, N = len(m["cells"]), F.num_basis(order)
    U, R = rd.f64(E * 5 * N).reshape(E, -1), rd.f64(E * 5 * N).reshape(E, -1)
    dt = rd.f64(1)[0]
    U1 = rd.f64(E * 5 * N).reshape(E, -1)
    assert (m["face_cells"][:, 1] >= 0).all() and (m["face_nodes_r"] != m["face_nodes"]).any()
    assert rel_err(R, oracle.residual(m, order, U, flux="roe", gamma=5 / 3)) < 1e-12
    assert rel_err(U1, oracle.ssprk3(m, order, U, dt, 1, flux="roe", gamma=5 / 3)) < 1e-11
    # ---- limiter, P2
    order = rd.i64()
    m, _ = rd.mesh()
    E, N = len(m["cells"]), F.num_basis(order)
    Uprev, U, A, B = (rd.f64(E * 5 * N).reshape(E, -1) for _ in range(4))
    expA = oracle.limiter_apply(order, U, oracle.limiter_bounds(m, order, Uprev), parts=3)
    bB = oracle.limiter_bounds(m, order, Uprev, only_avg=True)
    expB = oracle.limiter_apply(order, oracle.limiter_apply(order, U, bB, parts=1), bB, parts=2)
    assert rel_err(A, U) > 1e-3 and rel_err(B, U) > 1e-3  # the limiter acted
    assert rel_err(A, expA) < 1e-12 and rel_err(B, expB) < 1e-12
    # ---- implicit system, P1
    order = rd.i64()
    m, _ = rd.mesh()
    E, N = len(m["cells"]), F.num_basis(order)
    NB = 5 * N
    U = rd.f64(E * NB).reshape(E, NB)
    dt = rd.f64(1)[0]
    M, rhs, x, Ax = (rd.f64(E * NB).reshape(E, NB) for _ in range(4))
    iters, res = rd.i64(), rd.f64(1)[0]
    assert rel_err(M, oracle.mass(m, order)) < 1e-14
    diag, off, orhs = oracle.implicit_assemble(m, order, U)
    assert rel_err(rhs, orhs + M / dt * U) < 1e-12
    Ad = np.zeros((E * NB, E * NB))
    for e in range(E):
        Ad[e * NB:(e + 1) * NB, e * NB:(e + 1) * NB] += diag[e] + np.diag(M[e] / dt)
    for f, (c0, c1) in enumerate(m["face_cells"]):
        if c1 >= 0:
            Ad[c0 * NB:(c0 + 1) * NB, c1 * NB:(c1 + 1) * NB] += off[f, 0]
            Ad[c1 * NB:(c1 + 1) * NB, c0 * NB:(c0 + 1) * NB] += off[f, 1]
    assert iters > 0 and res < 1e-13
    assert rel_err(x.ravel(), np.linalg.solve(Ad, rhs.ravel())) < 1e-9
    assert rel_err(Ax.ravel(), Ad @ x.ravel()) < 1e-12
    assert rd.o == len(rd.b)
