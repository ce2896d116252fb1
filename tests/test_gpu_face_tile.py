"""GPU: the face-tile path (k_face_tile: traces, Riemann flux and face test of 16 faces of one
vertex-combination class on the matrix cores; k_elem_mfma adds the per-face contributions).

It serves the FP32 contractions at P2-P4 by default and FP64 with FVDG_FACE_TILE=all; the
switch is read when a context is created.  Bars (written per test): FP64 residual within 1e-12
of the oracle (the drop-in bar), FP32 residual within 2e-5 of the oracle and no further from it
than twice the FP32 path without face tiles (+1e-7), FP32 trajectory within 1e-6 (relative L2) of the FP64 one.
Meshes: a jittered periodic box (few classes), a box with Dirichlet and Pseudo3DZ faces
(boundary tiles), the reference's own box split (vertex orders from hashing: many classes) and
the reference's Fluent mesh case11.msh (unstructured, every orientation).
"""
import os
from pathlib import Path

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX, rel_err

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def make_solver(mode):
    old = os.environ.get("FVDG_FACE_TILE")
    os.environ["FVDG_FACE_TILE"] = mode
    try:
        return F.Solver(0)
    finally:
        if old is None:
            del os.environ["FVDG_FACE_TILE"]
        else:
            os.environ["FVDG_FACE_TILE"] = old


@pytest.fixture(scope="module")
def tile_all():
    s = make_solver("all")
    yield s
    s.close()


@pytest.fixture(scope="module")
def tile_off():
    s = make_solver("off")
    yield s
    s.close()


def l2_rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def meshes():
    return {
        "periodic": lambda: F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=5),
        "reference_split": lambda: F.Mesh.box((3, 2, 2), ub=(10, 10, 10), split=16),
    }


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [2, 3, 4])
@pytest.mark.parametrize("flux", ["lf", "roe"])
@pytest.mark.parametrize("mesh_kind", ["periodic", "reference_split"])
def test_face_tile_residual_against_oracle(tile_all, tile_off, oracle, order, flux, prec, mesh_kind):
    print(f"[case] face tile {prec} P{order} {flux} {mesh_kind}", flush=True)
    mesh = meshes()[mesh_kind]()
    periodic = mesh_kind == "periodic"
    params = VORTEX if periodic else [1.4, 5.0, 5.0, 5.0, 1.0, 0.5, 0.0, 0.0]
    U = mesh.project(order, "vortex", params)
    U *= 1 + 1e-2 * np.random.default_rng(order).standard_normal(U.shape)
    tile_all.set_scheme(order, flux=flux, precision=prec).set_mesh(mesh)
    kern = tile_all.stage_kernels()
    assert kern["flux"].startswith("k_face_tile"), kern
    tile_all.upload(U)
    m = mesh.export()
    UD = None
    if not periodic:
        _, xyz = tile_all.dirichlet_points()
        UD = np.array([[oracle.vortex_state(params, 0.0, x) for x in row] for row in xyz])
        tile_all.set_dirichlet_states(UD)
    Rt = tile_all.residual()
    Ro = oracle.residual(m, order, U, flux=flux, bc=None if periodic else mesh.boundary_types(), UD=UD)
    err = rel_err(Rt, Ro)
    if prec == "fp64":
        assert err < 1e-12, err
        return
    assert 1e-12 < err < 2e-5, err
    # against the FP32 path without face tiles (same splits, other summation order)
    tile_off.set_scheme(order, flux=flux, precision=prec).set_mesh(mesh)
    assert not tile_off.stage_kernels()["flux"].startswith("k_face_tile")
    tile_off.upload(U)
    if UD is not None:
        tile_off.set_dirichlet_states(UD)
    Rs = tile_off.residual()
    err_off = rel_err(Rs, Ro)
    print(f"fp32 vs oracle: face tiles {err:.2e}, without {err_off:.2e}", flush=True)
    assert err < 2 * err_off + 1e-7  # as accurate as the FP32 path without face tiles


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [2, 4])
def test_face_tile_parts_with_boundaries(tile_all, oracle, prec, order):
    """Dirichlet (x/y) and Pseudo3DZ (z) faces: boundary tiles, each residual part separately (FP32:
    the single parts take the per-face reference path, the whole residual the face tiles)."""
    print(f"[case] face tile parts {prec} P{order}", flush=True)
    mesh = F.Mesh.box((4, 4, 2), ub=(10, 10, 2.5), split=5)
    mesh.tag_pseudo3dz()
    bc = mesh.boundary_types()
    params = [1.4, 5.0, 4.0, 4.5, 1.0, 0.5, 0.0, 0.0]
    U = mesh.project(order, "vortex", params)
    tile_all.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
    tile_all.upload(U)
    _, xyz = tile_all.dirichlet_points()
    UD = np.array([[oracle.vortex_state(params, 0.0, x) for x in row] for row in xyz])
    tile_all.set_dirichlet_states(UD)
    m = mesh.export()
    tol = 1e-12 if prec == "fp64" else 2e-5
    for parts in (1, 2, 4, 6, 7):
        Rm = tile_all.residual(parts=parts)
        Ro = oracle.residual(m, order, U, flux="roe", bc=bc, UD=UD, parts=parts)
        assert rel_err(Rm, Ro) < tol, parts


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [2, 3])
def test_face_tile_flux_free_boundary_faces(tile_all, oracle, prec, order):
    """Boundary types without an explicit face flux (Neumann = 1, Outflow = 7: they contribute
    nothing) beside Dirichlet faces.  FP32 takes every face difference from the cell's mean flux,
    so a flux-free face must still carry that reference: a cell touching one is checked here."""
    print(f"[case] face tile flux-free faces {prec} P{order}", flush=True)
    mesh = F.Mesh.box((3, 3, 2), ub=(10, 10, 5), split=6, jitter=0.1, seed=7)
    bc = mesh.boundary_types()
    bnd = mesh.export()["face_cells"][:, 1] < 0
    idx = np.flatnonzero(bnd)
    bc[idx[0::3]] = 1
    bc[idx[1::3]] = 7
    mesh.set_boundary_types(bc)
    params = [1.4, 5.0, 5.0, 2.5, 1.0, 0.5, 0.0, 0.0]
    U = mesh.project(order, "vortex", params)
    tile_all.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
    tile_all.upload(U)
    _, xyz = tile_all.dirichlet_points()
    UD = np.array([[oracle.vortex_state(params, 0.0, x) for x in row] for row in xyz])
    tile_all.set_dirichlet_states(UD)
    Rm = tile_all.residual()
    Ro = oracle.residual(mesh.export(), order, U, flux="roe", bc=mesh.boundary_types(), UD=UD)
    assert rel_err(Rm, Ro) < (1e-12 if prec == "fp64" else 2e-5)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [2, 3])
def test_face_tile_fluent_mesh(tile_all, oracle, prec, order):
    """The reference's case11.msh: unstructured, faces in every vertex-combination class."""
    print(f"[case] face tile fluent {prec} P{order}", flush=True)
    mesh = F.Mesh.read_fluent(str(GOLDEN / "case11.msh"))
    for z, (bct, _name) in mesh.zones().items():
        if bct == 3:
            mesh.set_zone_boundary_type(z, F.BC_DIRICHLET)
    U = mesh.project(order, "const", [1.0, 0.2, 0.1, 0.0, 2.6])
    U *= 1 + 1e-2 * np.random.default_rng(order).standard_normal(U.shape)
    tile_all.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
    assert tile_all.stage_kernels()["flux"].startswith("k_face_tile")
    tile_all.upload(U)
    _, xyz = tile_all.dirichlet_points()
    UD = np.tile([1.0, 0.25, 0.05, 0.0, 2.7], xyz.shape[:2] + (1,))
    tile_all.set_dirichlet_states(UD)
    Rm = tile_all.residual()
    Ro = oracle.residual(mesh.export(), order, U, flux="roe", bc=mesh.boundary_types(), UD=UD.reshape(-1, 5))
    assert rel_err(Rm, Ro) < (1e-12 if prec == "fp64" else 2e-5)


@pytest.mark.parametrize("order", [3, 4])
def test_face_tile_trajectory_tracks_fp64(tile_all, order):
    """40 SSP-RK3 steps: FP32 face tiles within 1e-6 (relative L2) of the FP64 path, conserved means."""
    print(f"[case] face tile trajectory P{order}", flush=True)
    mesh = F.Mesh.box(4, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=3)
    U = mesh.project(order, "vortex", VORTEX)
    out = {}
    for prec in ("fp64", "fp32"):
        tile_all.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
        tile_all.upload(U)
        s0 = tile_all.conservation_sums()
        tile_all.ssprk3(0.01, 40)
        out[prec] = tile_all.download()
        s1 = tile_all.conservation_sums()
        # relative to the largest conserved integral (the rho w integral of this flow is zero);
        # FP32 sums the four face contributions of a cell in FP32
        drift = np.abs(np.asarray(s1) - np.asarray(s0)) / np.abs(s0).max()
        assert drift.max() < (1e-13 if prec == "fp64" else 1e-11), (prec, drift)
    assert l2_rel(out["fp32"], out["fp64"]) < 1e-6
