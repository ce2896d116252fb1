"""Flow diagnostics of the device state (fvdg_flow_diagnostics): volume, mass, kinetic energy,
enstrophy and total energy integrals, and the point extremes of density, pressure and Mach number.

Pinned by closed forms -- a solid-body rotation u = s (-(y - 1/2), x - 1/2, 0) on the unit cube,
exactly representable from P1 up, has kinetic energy s^2 / 12 and enstrophy 2 s^2; the
compressible Taylor-Green vortex has E_k = 1/8 and epsilon = 3/8 (per rho_0 |Omega|), approached
as the mesh is refined -- and by an independent evaluation in numpy (the oracle's basis and
geometry, numpy's matrix inverse) at 1e-12.
"""
import numpy as np
import pytest

import fve_dg_amd as F
from conftest import rel_err

pytestmark = pytest.mark.gpu
GAMMA = 1.4
TGV = [GAMMA, 1.0, 1.0, 0.1]  # gamma, rho0, V0, Mach


def project_linear(oracle, m, order, fn):
    """diagonal L2 projection of a pointwise state fn(x) -> (5,) (exact for linear fields)"""
    pts, w = oracle.quadrature(order, 0)
    og = oracle.geometry(m)
    Phi = np.array([oracle.basis(order, *x)[0] for x in pts])  # Q x N
    diag = (w[:, None] * Phi ** 2).sum(0)
    U = np.zeros((len(m["cells"]), 5 * Phi.shape[1]))
    for e, cell in enumerate(m["cells"]):
        p0 = m["points"][cell[0]]
        X = p0 + pts @ og["A"][e]  # x(xi) = p0 + A^T xi, rows of A = edges
        vals = np.array([fn(x) for x in X])  # Q x 5
        U[e] = ((w[:, None, None] * Phi[:, :, None] * vals[:, None, :]).sum(0) / diag[:, None]).ravel()
    return U


def numpy_diagnostics(oracle, m, order, U, gamma=GAMMA):
    pts, w = oracle.quadrature(order, 0)
    og = oracle.geometry(m)
    B = [oracle.basis(order, *x) for x in pts]
    Phi = np.array([b[0] for b in B])  # Q x N
    Gr = np.array([b[1] for b in B])  # Q x N x 3
    N = Phi.shape[1]
    C = U.reshape(len(U), N, 5)
    u = np.einsum("qn,enk->eqk", Phi, C)
    gref = np.einsum("qnd,enk->eqdk", Gr, C)
    Ainv = np.linalg.inv(og["A"])
    gx = np.einsum("eij,eqjk->eqik", Ainv, gref)  # grad_x = A^{-1} grad_xi
    wd = w[None, :] * og["det"][:, None]
    rho = u[..., 0]
    vel = u[..., 1:4] / rho[..., None]
    du = (gx[..., :, 1:4] - vel[..., None, :] * gx[..., :, 0:1]) / rho[..., None, None]  # [e,q,d,i] = d u_i / d x_d
    om = np.stack([du[..., 1, 2] - du[..., 2, 1], du[..., 2, 0] - du[..., 0, 2], du[..., 0, 1] - du[..., 1, 0]], -1)
    v2 = (vel ** 2).sum(-1)
    p = (gamma - 1) * (u[..., 4] - 0.5 * rho * v2)
    return dict(volume=wd.sum(), mass=(wd * rho).sum(), kinetic_energy=(wd * 0.5 * rho * v2).sum(),
                enstrophy=(wd * 0.5 * rho * (om ** 2).sum(-1)).sum(), total_energy=(wd * u[..., 4]).sum(),
                min_rho=rho.min(), min_p=p.min(), max_mach=(np.sqrt(v2) / np.sqrt(gamma * p / rho)).max())


@pytest.mark.parametrize("order", [1, 3])
def test_solid_body_rotation_closed_forms(solver, oracle, order):
    print(f"[case] diagnostics rotation P{order}", flush=True)
    s = 0.7
    mesh = F.Mesh.box(3, periodic=0)
    m = mesh.export()
    U = project_linear(oracle, m, order, lambda x: np.array([1.0, -s * (x[1] - 0.5), s * (x[0] - 0.5), 0.0, 10.0]))
    solver.set_scheme(order).set_mesh(mesh)
    solver.upload(U)
    d = solver.flow_diagnostics()
    assert abs(d["volume"] - 1) < 1e-13 and abs(d["mass"] - 1) < 1e-13 and abs(d["total_energy"] - 10) < 1e-12
    assert abs(d["kinetic_energy"] - s * s / 12) < 1e-13
    assert abs(d["enstrophy"] - 2 * s * s) < 1e-12
    assert abs(d["min_rho"] - 1) < 1e-13
    ref = numpy_diagnostics(oracle, m, order, U)
    for k, v in ref.items():
        assert abs(d[k] - v) <= 1e-12 * max(abs(v), 1.0), k


def test_taylor_green_vortex_closed_forms_and_numpy(solver, oracle):
    """P3: E_k converges at O(h^8) (the projection's error is orthogonal to the field) and the
    enstrophy, a broken-gradient quantity, at O(h^3) (element-boundary terms)"""
    print("[case] diagnostics TGV", flush=True)
    L = 2 * np.pi
    vol = L ** 3
    out = []
    for n in (4, 8):
        mesh = F.Mesh.box(n, ub=(L, L, L), periodic=7)
        U = mesh.project(3, "tgv", TGV)
        solver.set_scheme(3).set_mesh(mesh)
        solver.upload(U)
        d = solver.flow_diagnostics()
        assert abs(d["volume"] - vol) < 1e-12 * vol
        out.append((abs(d["kinetic_energy"] / vol - 0.125), abs(d["enstrophy"] / vol - 0.375)))
        if n == 4:
            ref = numpy_diagnostics(oracle, mesh.export(), 3, U)
            for k, v in ref.items():
                assert abs(d[k] - v) <= 1e-12 * max(abs(v), 1.0), k
    (ek4, en4), (ek8, en8) = out
    assert ek8 < 1e-7 and en8 < 1e-3
    assert ek8 < ek4 / 64 and en8 < en4 / 6
