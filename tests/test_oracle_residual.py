"""Pins of the oracle's residual and time stepping against what the mathematics fixes.

* divergence identity: for a state whose flux is a polynomial the quadrature rules integrate
  exactly, with exact Dirichlet data (or the Pseudo3DZ mirror of a flow with w = 0), the DG
  residual equals int_e (div F) phi_j dV -- computed here from a closed-form div F.  A dropped
  term, a wrong sign on one side of a face, a wrong normal or a transposed inverse Jacobian
  each break it;
* P0 is the first-order finite-volume scheme sum_f +-A_f F^(U_L, U_R; n_f);
* global conservation on periodic meshes (sum_e R_e,0 = 0, phi_0 = 1);
* SSP-RK3 equals the Shu-Osher three-stage form composed from residual and mass;
* the vortex field is an exact solution of the Euler equations (finite-difference check);
* the threaded (atomic) oracle equals the serial one.
"""
from pathlib import Path

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX, rel_err

GAMMA = 1.4
GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_meshes.npz"


# ----------------------------------------------------------------- helpers
class LinearState:
    """rho = 1, momentum a + B x, total energy e0 + g.x (all in the P1 space)."""

    def __init__(self, seed, planar=False):
        r = np.random.default_rng(seed)
        self.a = r.uniform(-0.3, 0.3, 3)
        self.B = r.uniform(-0.05, 0.05, (3, 3))
        self.e0 = 3.0
        self.g = r.uniform(-0.05, 0.05, 3)
        if planar:  # a 2D flow: no w, nothing depends on z
            self.a[2] = 0.0
            self.B[2, :] = 0.0
            self.B[:, 2] = 0.0
            self.g[2] = 0.0

    def state(self, x):
        x = np.atleast_2d(x)
        m = self.a + x @ self.B.T
        E = self.e0 + x @ self.g
        return np.column_stack([np.ones(len(x)), m, E])

    def div_flux(self, x):
        """closed-form div F of the Euler flux (rho = 1, so u = m)"""
        x = np.atleast_2d(x)
        u = self.a + x @ self.B.T
        E = self.e0 + x @ self.g
        p = (GAMMA - 1) * (E - 0.5 * (u * u).sum(1))
        gp = (GAMMA - 1) * (self.g - u @ self.B)  # grad p = (gamma-1)(g - B^T u)
        tr = np.trace(self.B)
        out = np.zeros((len(x), 5))
        out[:, 0] = tr
        out[:, 1:4] = u @ self.B.T + u * tr[None] + gp
        out[:, 4] = ((self.g + gp) * u).sum(1) + (E + p) * tr
        return out


def cell_points(m, xi):
    """physical points x = X0 + sum_d xi_d (X_d - X0) of reference points xi, per cell"""
    X = m["points"][m["cells"]]  # (E, 4, 3)
    lam = np.column_stack([1 - xi.sum(1), xi])  # (Q, 4)
    return np.einsum("qa,ead->eqd", lam, X)


def project(m, order, fn):
    """L2 projection with the volume rule (exact for polynomial fields of degree <= order)"""
    vp, vw, _, _ = F.quadrature(order)
    N = F.num_basis(order)
    phi = np.array([F.basis(order, *p)[0] for p in vp])  # (Q, N)
    mass_hat = (phi * phi * vw[:, None]).sum(0)
    X = cell_points(m, vp)
    E = len(m["cells"])
    vals = fn(X.reshape(-1, 3)).reshape(E, len(vw), 5)
    c = np.einsum("q,qk,eqv->ekv", vw, phi, vals) / mass_hat[None, :, None]
    return c.reshape(E, 5 * N)


def divergence_rhs(m, order, det, fn):
    """int_e (div F) phi_j dV with the volume rule"""
    vp, vw, _, _ = F.quadrature(order)
    phi = np.array([F.basis(order, *p)[0] for p in vp])
    X = cell_points(m, vp)
    E = len(m["cells"])
    d = fn(X.reshape(-1, 3)).reshape(E, len(vw), 5)
    return np.einsum("q,qk,eqv,e->ekv", vw, phi, d, det).reshape(E, -1)


def dirichlet_data(m, order, og, bc, fn):
    """exact states at the face points of the Dirichlet faces, in face order"""
    _, _, fp, _ = F.quadrature(order)
    rows = []
    for f in np.nonzero((m["face_cells"][:, 1] < 0) & (bc == 0))[0]:
        c = m["face_cells"][f, 0]
        vt = og["vert"][f, 0]
        X = m["points"][m["cells"][c]]
        for u, v in fp:
            lam = np.zeros(4)
            lam[vt[0]] += 1 - u - v
            lam[vt[1]] += u
            lam[vt[2]] += v
            rows.append(fn(lam @ X)[0])
    return np.array(rows)


def face_scale(oracle, m, og, U):
    """size of one face term, area x |F|: the face rule's integration error scales with it"""
    return og["area"].max() * max(np.abs(oracle.flux(u)).max() for u in U[:, :5])


def meshes():
    d = np.load(GOLDEN)
    ref = {n: {k: d[f"{n}_{k}"] for k in ("points", "cells", "face_nodes", "face_cells")} for n in ("box6", "box5")}
    return {
        "ref_box6": F.Mesh.from_arrays(**ref["box6"]).export(),
        "ref_box5": F.Mesh.from_arrays(**ref["box5"]).export(),
        "jittered": F.Mesh.box(3, ub=(1, 1, 1), periodic=0, jitter=0.15, seed=4).export(),
    }


MESHES = meshes()


# ----------------------------------------------------------------- divergence identity
@pytest.mark.parametrize("mesh", sorted(MESHES))
@pytest.mark.parametrize("order,flux", [(1, "lf"), (2, "roe"), (3, "lf"), (4, "roe")])
def test_divergence_identity_dirichlet(oracle, mesh, order, flux):
    m = MESHES[mesh]
    og = oracle.geometry(m)
    st = LinearState(seed=order)
    U = project(m, order, st.state)
    bc = np.zeros(len(m["face_nodes"]), np.uint8)
    UD = dirichlet_data(m, order, og, bc, st.state)
    R = oracle.residual(m, order, U, flux=flux, bc=bc, UD=UD)
    Rx = divergence_rhs(m, order, og["det"], st.div_flux)
    # P1 uses exact rules (error at rounding level of R); the P2-P4 face rule integrates its
    # polynomials to ~1e-10 only, an error relative to the size of a single face term
    if order == 1:
        assert np.abs(R - Rx).max() <= 1e-12 * np.abs(Rx).max()
    else:
        assert np.abs(R - Rx).max() <= 1e-9 * face_scale(oracle, m, og, U)


@pytest.mark.parametrize("order", [1, 3])
def test_divergence_identity_pseudo3dz(oracle, order):
    """z faces tagged Pseudo3DZ (mirror of rho w), other boundary faces Dirichlet"""
    mesh = F.Mesh.box((3, 3, 2), ub=(1, 1, 0.5), periodic=0, jitter=0.1, seed=2)
    mesh.tag_pseudo3dz()
    m = mesh.export()
    bc = mesh.boundary_types()
    bnd = m["face_cells"][:, 1] < 0
    assert (bc[bnd] == 5).sum() == 2 * 2 * 9 and (bc[bnd] == 0).sum() == 2 * 2 * 2 * 6
    og = oracle.geometry(m)
    st = LinearState(seed=10 + order, planar=True)
    U = project(m, order, st.state)
    UD = dirichlet_data(m, order, og, bc, st.state)
    R = oracle.residual(m, order, U, flux="roe", bc=bc, UD=UD)
    Rx = divergence_rhs(m, order, og["det"], st.div_flux)
    scale = np.abs(Rx).max() if order == 1 else face_scale(oracle, m, og, U)
    assert np.abs(R - Rx).max() <= (1e-12 if order == 1 else 1e-9) * scale


def test_divergence_identity_detects_a_planted_sign_error(oracle):
    """The identity is sensitive: wrong Dirichlet data (half the exact state) breaks it
    by orders of magnitude beyond the tolerance."""
    m = MESHES["jittered"]
    og = oracle.geometry(m)
    st = LinearState(seed=1)
    U = project(m, 2, st.state)
    bc = np.zeros(len(m["face_nodes"]), np.uint8)
    UD = dirichlet_data(m, 2, og, bc, st.state)
    Rx = divergence_rhs(m, 2, og["det"], st.div_flux)
    R_bad = oracle.residual(m, 2, U, flux="lf", bc=bc, UD=0.5 * UD)
    assert np.abs(R_bad - Rx).max() > 1e-3 * np.abs(Rx).max()


# ----------------------------------------------------------------- P0 = finite volumes
@pytest.mark.parametrize("flux", ["lf", "roe"])
def test_p0_is_first_order_finite_volume(oracle, flux):
    mesh = F.Mesh.box(3, ub=(1, 1, 1), periodic=7, jitter=0.2, seed=8)
    m = mesh.export()
    og = oracle.geometry(m)
    rng = np.random.default_rng(0)
    E = len(m["cells"])
    rho = rng.uniform(0.5, 1.5, E)
    vel = rng.uniform(-0.5, 0.5, (E, 3))
    p = rng.uniform(0.5, 1.5, E)
    U = np.column_stack([rho, rho[:, None] * vel, p / (GAMMA - 1) + 0.5 * rho * (vel ** 2).sum(1)])
    R = oracle.residual(m, 0, U, flux=flux)
    Rx = np.zeros_like(U)
    for f, (c0, c1) in enumerate(m["face_cells"]):
        Fh = oracle.numflux(flux, U[c0], U[c1], og["normal"][f])
        Rx[c0] += og["area"][f] * Fh
        Rx[c1] -= og["area"][f] * Fh
    assert rel_err(R, Rx) < 1e-13
    M = oracle.mass(m, 0)
    assert np.allclose(M[:, 0], og["det"] / 6, rtol=1e-14)  # P0 mass = cell volume


# ----------------------------------------------------------------- conservation and freestream
@pytest.mark.parametrize("order,flux", [(1, "roe"), (3, "lf"), (4, "roe")])
def test_global_conservation_periodic(oracle, order, flux):
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=1)
    m = mesh.export()
    U = oracle.project_vortex(m, order, VORTEX)
    R = oracle.residual(m, order, U, flux=flux)
    s = R[:, :5].sum(0)  # phi_0 = 1: the cell means' residuals sum to zero
    assert np.abs(s).max() < 1e-12 * np.abs(R[:, :5]).sum(0).max()


@pytest.mark.parametrize("order", [0, 1, 2])
def test_freestream_preserved(oracle, order):
    mesh = F.Mesh.box(3, ub=(1, 1, 1), periodic=7, jitter=0.2, seed=6)
    m = mesh.export()
    st = LinearState(seed=0)
    st.B[:] = 0
    st.g[:] = 0
    U = project(m, order, st.state)
    R = oracle.residual(m, order, U, flux="roe")
    # P0/P1 rules exact; the P2 face rule carries its ~1e-10 integration error
    assert np.abs(R).max() < (1e-13 if order < 2 else 1e-9)


# ----------------------------------------------------------------- time stepping
def test_ssprk3_is_shu_osher(oracle):
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=3)
    m = mesh.export()
    U0 = oracle.project_vortex(m, 2, VORTEX)
    M = oracle.mass(m, 2)
    dt = 0.02
    L = lambda U: -oracle.residual(m, 2, U, flux="roe") / M
    U1 = U0 + dt * L(U0)
    U2 = 0.75 * U0 + 0.25 * (U1 + dt * L(U1))
    U3 = U0 / 3 + 2 / 3 * (U2 + dt * L(U2))
    got = oracle.ssprk3(m, 2, U0, dt, 1, flux="roe")
    assert rel_err(got, U3) < 1e-14


def test_threaded_oracle_equals_serial(oracle):
    mesh = F.Mesh.box(4, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=3)
    m = mesh.export()
    U = oracle.project_vortex(m, 3, VORTEX)
    a = oracle.residual(m, 3, U, flux="lf", threads=1)
    b = oracle.residual(m, 3, U, flux="lf", threads=4)
    assert rel_err(a, b) < 1e-13


# ----------------------------------------------------------------- the vortex field
def test_vortex_is_an_exact_euler_solution(oracle):
    """dU/dt + div F(U) = 0 for the analytic field (centred differences, h = 1e-3)."""
    rng = np.random.default_rng(1)
    P = np.array([1.4, 5.0, 5.0, 5.0, 1.0, 1.0, 0.0, 0.0])  # no periodic wrap
    h = 1e-3
    worst = 0.0
    for _ in range(20):
        x = np.array([rng.uniform(3.5, 6.5), rng.uniform(3.5, 6.5), 0.0])
        t = rng.uniform(0, 0.5)
        dUdt = (oracle.vortex_state(P, t + h, x) - oracle.vortex_state(P, t - h, x)) / (2 * h)
        div = np.zeros(5)
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            div += (oracle.flux(oracle.vortex_state(P, t, x + e))[:, d]
                    - oracle.flux(oracle.vortex_state(P, t, x - e))[:, d]) / (2 * h)
        worst = max(worst, np.abs(dUdt + div).max())
    assert worst < 1e-4  # O(h^2) truncation; a wrong exponent or sign gives O(1)
    # far field: free stream rho = p = 1, velocity (u_inf, v_inf)
    Uf = oracle.vortex_state(P, 0.0, np.array([5.0 + 40, 5.0, 0.0]))
    assert np.allclose(Uf, [1, 1, 1, 0, 1 / 0.4 + 1.0], atol=1e-12)


def test_library_projection_matches_oracle(oracle):
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=12)
    m = mesh.export()
    for order in (0, 2, 4):
        a = mesh.project(order, "vortex", VORTEX)
        b = oracle.project_vortex(m, order, VORTEX)
        assert rel_err(a, b) < 1e-14
