"""Fluent .msh ingestion (the reference's data/case11.msh and data/case46.msh, kept as fixtures in
tests/golden).

Checked against an independent reading of the same file in Python: the domain volume (sum of
the cells' volumes from their face lists by the divergence theorem) equals the sum of the
tetrahedra, every boundary polygon becomes (nodes - 2) boundary triangles of its zone, and a
uniform state is preserved by the DG residual on the resulting unstructured mesh (oracle).
"""
import re
from pathlib import Path

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import rel_err

CASE11 = Path(__file__).resolve().parent / "golden" / "case11.msh"
CASE46 = Path(__file__).resolve().parent / "golden" / "case46.msh"  # the reference's data/case46.msh


def parse_fluent(path):
    """minimal independent reader: nodes, and faces as (nodes, c0, c1, zone)"""
    text = Path(path).read_text()
    nodes = {}
    faces = []
    for m in re.finditer(r"\((10|13)\s*\(([0-9a-fA-F\s]+)\)\s*\(", text):
        sec = m.group(1)
        head = [int(x, 16) for x in m.group(2).split()]
        body = text[m.end(): text.index(")", m.end())].split()
        if sec == "10":
            first, last = head[1], head[2]
            xyz = np.array(body, dtype=float).reshape(-1, 3)
            for i, p in zip(range(first, last + 1), xyz):
                nodes[i] = p
        else:
            zone, first, last, ftype = head[0], head[1], head[2], head[4]
            vals = [int(x, 16) for x in body]
            k = 0
            for _ in range(first, last + 1):
                n = vals[k] if ftype == 0 else ftype
                k += 1 if ftype == 0 else 0
                faces.append((vals[k:k + n], vals[k + n], vals[k + n + 1], zone))
                k += n + 2
    return nodes, faces


def test_case11_reads_and_matches_an_independent_parse():
    mesh = F.Mesh.read_fluent(CASE11)
    zones = mesh.zones()
    assert {v[1] for v in zones.values()} >= {"SYM1", "SYM2", "DRIVWALL", "WALL"}
    assert zones[9][0] == 3  # wall
    check_against_parse(mesh, CASE11)


def test_case46_reads_and_matches_an_independent_parse():
    """data/case46.msh: 8,000 hexahedra (each split into 12 tets around its centroid), zones BOT,
    INLET (velocity inlet, type 10), OUTLET (pressure outlet, 5), SYM, TOP"""
    mesh = F.Mesh.read_fluent(CASE46)
    assert mesh.sizes[1] == 96000
    zones = mesh.zones()
    assert {v[1] for v in zones.values()} == {"FLUID", "int_FLUID", "BOT", "INLET", "OUTLET", "SYM", "TOP"}
    assert {v[0] for v in zones.values() if v[1] in ("INLET", "OUTLET")} == {10, 5}
    check_against_parse(mesh, CASE46)


def check_against_parse(mesh, path):
    nodes, faces = parse_fluent(path)
    # domain volume: sum over cells of (1/3) sum_faces (x_f . n_f) A_f with outward normals,
    # i.e. over all faces: (1/3) x_f . (n A)_f counted for c0 with the face orientation and
    # for c1 opposite -- interior faces cancel, leaving the boundary integral
    vol = 0.0
    for nds, c0, c1, _z in faces:
        P = np.array([nodes[i] for i in nds])
        area = np.zeros(3)
        for i in range(1, len(P) - 1):
            area += 0.5 * np.cross(P[i] - P[0], P[i + 1] - P[0])
        xc = P.mean(0)
        if c1 == 0:
            vol += np.dot(xc, area) / 3.0  # boundary faces of this file point out of their cell
        elif c0 == 0:
            vol -= np.dot(xc, area) / 3.0
    g = mesh.geometry()
    assert abs(abs(vol) - g["det"].sum() / 6) < 1e-12 * abs(vol)
    assert (g["det"] > 0).all()
    # boundary triangles: each boundary polygon of zone z gives (nodes - 2) triangles of zone z
    a = mesh.export()
    bnd = a["face_cells"][:, 1] < 0
    fz = mesh.face_zones()
    expect = {}
    for nds, c0, c1, z in faces:
        if c0 == 0 or c1 == 0:
            expect[z] = expect.get(z, 0) + len(nds) - 2
    got = dict(zip(*np.unique(fz[bnd], return_counts=True)))
    assert got == expect


def test_zone_boundary_types_and_freestream(oracle):
    mesh = F.Mesh.read_fluent(CASE11)
    for z, (bct, _name) in mesh.zones().items():
        if bct != 2 and bct != 0:
            mesh.set_zone_boundary_type(z, F.BC_DIRICHLET)
    bc = mesh.boundary_types()
    a = mesh.export()
    bnd = a["face_cells"][:, 1] < 0
    assert (bc[bnd] == F.BC_DIRICHLET).all()
    Uc = np.array([1.2, 0.3, -0.2, 0.0, 3.1])
    for order in (1, 2):
        U = mesh.project(order, "const", Uc)
        _, _, fp, _ = F.quadrature(order)
        UD = np.tile(Uc, (int(bnd.sum()) * len(fp), 1))
        R = oracle.residual(a, order, U, flux="roe", bc=bc, UD=UD)
        scale = np.abs(oracle.flux(Uc)).max() * oracle.geometry(a)["area"].max()
        assert np.abs(R).max() < (1e-12 if order == 1 else 1e-9) * scale
    with pytest.raises(F.FvdgError):
        mesh.set_zone_boundary_type(8, F.BC_DIRICHLET)  # the interior zone has no boundary face


def test_missing_file_is_an_error():
    with pytest.raises(F.FvdgError):
        F.Mesh.read_fluent(CASE11.with_name("no_such_file.msh"))


@pytest.mark.gpu
@pytest.mark.parametrize("order,prec", [(1, "fp64"), (2, "fp64"), (3, "fp64"), (3, "fp32")])
def test_fluent_mesh_residual_matches_oracle(solver, oracle, order, prec):
    """case11.msh (unstructured, any vertex combination on any face side) with Dirichlet zones:
    FP64 to the parity bar, FP32 contractions to 2e-5."""
    print(f"[case] case11.msh residual P{order} {prec}", flush=True)
    mesh = F.Mesh.read_fluent(CASE11)
    for z, (bct, _name) in mesh.zones().items():
        if bct == 3:
            mesh.set_zone_boundary_type(z, F.BC_DIRICHLET)
    a = mesh.export()
    U = mesh.project(order, "const", [1.0, 0.2, 0.1, 0.0, 2.6])
    U *= 1 + 1e-2 * np.random.default_rng(0).standard_normal(U.shape)
    solver.set_scheme(order, flux="roe", precision=prec).set_mesh(mesh)
    solver.upload(U)
    _, xyz = solver.dirichlet_points()
    UD = np.tile([1.0, 0.25, 0.05, 0.0, 2.7], xyz.shape[:2] + (1,))
    solver.set_dirichlet_states(UD)
    Rg = solver.residual()
    Ro = oracle.residual(a, order, U, flux="roe", bc=mesh.boundary_types(), UD=UD.reshape(-1, 5))
    assert rel_err(Rg, Ro) < (1e-12 if prec == "fp64" else 2e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("order,prec", [(1, "fp64"), (2, "fp64"), (3, "fp32")])
def test_case46_residual_matches_oracle(solver, oracle, order, prec):
    """case46.msh (96,000 tets from 8,000 hexahedra) with Dirichlet data on its wall zones (BOT,
    SYM, TOP) and the inlet/outlet zones contributing nothing, as in the reference: FP64 to the
    parity bar, FP32 contractions to 2e-5."""
    print(f"[case] case46.msh residual P{order} {prec}", flush=True)
    mesh = F.Mesh.read_fluent(CASE46)
    for z, (bct, _name) in mesh.zones().items():
        if bct == 3:
            mesh.set_zone_boundary_type(z, F.BC_DIRICHLET)
    a = mesh.export()
    U = mesh.project(order, "const", [1.0, 0.2, 0.1, 0.0, 2.6])
    U *= 1 + 1e-2 * np.random.default_rng(1).standard_normal(U.shape)
    solver.set_scheme(order, flux="lf", precision=prec).set_mesh(mesh)
    solver.upload(U)
    _, xyz = solver.dirichlet_points()
    UD = np.tile([1.0, 0.25, 0.05, 0.0, 2.7], xyz.shape[:2] + (1,))
    solver.set_dirichlet_states(UD)
    Rg = solver.residual()
    Ro = oracle.residual(a, order, U, flux="lf", bc=mesh.boundary_types(), UD=UD.reshape(-1, 5),
                         threads=oracle.max_threads())
    assert rel_err(Rg, Ro) < (1e-12 if prec == "fp64" else 2e-5)
