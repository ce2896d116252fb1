"""Mesh construction and geometry: counts, periodic pairing, geometry identities,
agreement of the library's and the oracle's geometry, and the reference's own meshes."""
from pathlib import Path

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import rel_err

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_meshes.npz"


@pytest.mark.parametrize("split", [5, 6])
@pytest.mark.parametrize("periodic", [0, 7])
def test_box_counts(split, periodic):
    n = (2, 4, 2)
    m = F.Mesh.box(n, ub=(1, 2, 1), split=split, periodic=periodic)
    npnt, nc, nf = m.sizes
    hexes = n[0] * n[1] * n[2]
    assert npnt == (n[0] + 1) * (n[1] + 1) * (n[2] + 1)
    assert nc == split * hexes
    a = m.export()
    nb = (a["face_cells"][:, 1] < 0).sum()
    # every tet has 4 faces; interior faces are shared by 2 tets
    assert 4 * nc == 2 * (nf - nb) + nb
    if periodic == 7:
        assert nb == 0
        # one merged face per pair of opposite boundary triangles (2 per boundary square)
        per = (a["face_nodes_r"] != a["face_nodes"]).any(1).sum()
        assert per == 2 * (n[0] * n[1] + n[1] * n[2] + n[0] * n[2])
    else:
        assert nb == 2 * 2 * (n[0] * n[1] + n[1] * n[2] + n[0] * n[2])


def test_geometry_identities_and_oracle_agreement(oracle):
    m = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=5)
    a = m.export()
    g = m.geometry()
    og = oracle.geometry(a)
    # same formulas; the two compilers may contract different products into FMAs (1 ulp)
    assert rel_err(g["det"], og["det"]) < 1e-14
    assert rel_err(g["normal"], og["normal"]) < 1e-14
    assert rel_err(g["area"], og["area"]) < 1e-14
    # A^{-1} A = I, |det A| = 6 V > 0
    for e in range(len(a["cells"])):
        assert np.abs(g["inv_jac"][e] @ og["A"][e] - np.eye(3)).max() < 1e-12
    assert (g["det"] > 0).all()
    assert abs(g["det"].sum() / 6 - 1000.0) < 1e-9  # the cell volumes tile the box
    # closure: sum over a tet's faces of the outward area vector is zero
    closure = np.zeros((len(a["cells"]), 3))
    for f, (c0, c1) in enumerate(a["face_cells"]):
        closure[c0] += g["area"][f] * g["normal"][f]
        if c1 >= 0:
            closure[c1] -= g["area"][f] * g["normal"][f]
    assert np.abs(closure).max() < 1e-12 * g["area"].max() * 10


def test_face_points_coincide_from_both_sides(oracle):
    """A face quadrature point mapped through either neighbour's reference coordinates
    lands on the same physical point (modulo the period across periodic faces)."""
    m = F.Mesh.box(2, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=9)
    a = m.export()
    og = oracle.geometry(a)
    _, _, fp, _ = F.quadrature(3)
    P = a["points"]
    for f in range(len(a["face_nodes"])):
        pts = []
        for side in (0, 1):
            c = a["face_cells"][f, side]
            vt = og["vert"][f, side]
            X = P[a["cells"][c]]
            for u, v in fp:
                lam = np.zeros(4)
                lam[vt[0]] += 1 - u - v
                lam[vt[1]] += u
                lam[vt[2]] += v
                pts.append(lam @ X)
        pts = np.array(pts).reshape(2, -1, 3)
        d = pts[0] - pts[1]
        d -= 10.0 * np.round(d / 10.0)  # periodic images
        assert np.abs(d).max() < 1e-12


def test_jitter_keeps_periodic_images_consistent():
    m = F.Mesh.box(4, ub=(1, 1, 1), periodic=7, jitter=0.2, seed=3)
    a = m.export()
    P = a["points"]
    # points on x = 0 and x = 1 (before jitter) are images: equal y, z and x offset by 1
    n = 5
    idx = lambda i, j, k: i + (j + k * n) * n
    for j in range(n):
        for k in range(n):
            lo, hi = P[idx(0, j, k)], P[idx(4, j, k)]
            assert np.allclose(hi - lo, [1.0, 0, 0], atol=1e-14)
    assert np.abs(P[idx(2, 2, 2)] - [0.5, 0.5, 0.5]).max() > 0  # interior points moved


def test_non_periodic_jitter_keeps_boundary_planes():
    m = F.Mesh.box(3, ub=(1, 1, 1), periodic=0, jitter=0.2, seed=3)
    P = m.export()["points"]
    on = np.isclose(P, 0.0) | np.isclose(P, 1.0)
    assert on.sum() > 0
    g = m.geometry()
    assert abs(g["det"].sum() / 6 - 1.0) < 1e-12


def test_invalid_meshes_are_rejected():
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    cells = np.array([[0, 1, 2, 3]])
    faces = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    fc = np.array([[0, -1]] * 4)
    F.Mesh.from_arrays(pts, cells, faces, fc)  # valid
    with pytest.raises(F.FvdgError):
        F.Mesh.from_arrays(pts, cells, faces, np.array([[1, -1]] * 4))  # face cell out of range
    with pytest.raises(F.FvdgError):
        bad = faces.copy()
        bad[0] = [0, 1, 2]  # face 0 duplicates face 3: cell face listed twice
        F.Mesh.from_arrays(pts, cells, bad, fc)
    with pytest.raises(F.FvdgError):
        flat = pts.copy()
        flat[3] = [1, 1, 0]  # degenerate (coplanar) tetrahedron
        F.Mesh.from_arrays(flat, cells, faces, fc)
    with pytest.raises(F.FvdgError):
        F.Mesh.box(3, split=5, periodic=1)  # odd count along a periodic axis with the 5-split


def test_reference_meshes_load(oracle):
    d = np.load(GOLDEN)
    for name, ncell in (("box6", 108), ("box5", 60)):
        a = {k: d[f"{name}_{k}"] for k in ("points", "cells", "face_nodes", "face_cells")}
        m = F.Mesh.from_arrays(**a)
        assert m.sizes[1] == ncell
        g, og = m.geometry(), oracle.geometry(a)
        assert rel_err(g["normal"], og["normal"]) < 1e-14 and rel_err(g["det"], og["det"]) < 1e-14
        assert abs(g["det"].sum() / 6 - 250.0) < 1e-9


def test_reference_mesh_fixture_matches_reference_build(ref):
    """The committed fixture is what the reference's mesh code produces."""
    d = np.load(GOLDEN)
    m = ref.box_mesh((3, 3, 2), (0, 0, 0), (10, 10, 2.5), 6)
    for k in ("points", "cells", "face_nodes", "face_cells"):
        assert np.array_equal(m[k], d[f"box6_{k}"])


def _export4(m):
    a = m.export()
    return {k: a[k] for k in ("points", "cells", "face_nodes", "face_cells")}


def test_reference_split_reproduces_the_fixture():
    """split 16 / 15 (OrthHexMesh + split_hex6_scan / split_hex5_scan + rebuild_cell_topology)
    gives the reference's mesh exactly: the committed fixture, built by the reference's code."""
    d = np.load(GOLDEN)
    for name, n, split in (("box6", (3, 3, 2), 16), ("box5", (2, 3, 2), 15)):
        m = _export4(F.Mesh.box(n, ub=(10, 10, 2.5), split=split))
        for k in ("points", "cells", "face_nodes", "face_cells"):
            assert np.array_equal(m[k], d[f"{name}_{k}"]), (name, k)


@pytest.mark.parametrize("n,ub", [((1, 1, 1), (1, 1, 1)), ((2, 2, 2), (1, 1, 1)), ((4, 4, 1), (10, 10, 2.5)),
                                  ((5, 3, 2), (2, 1, 3)), ((8, 8, 8), (10, 10, 10)), ((16, 16, 1), (10, 10, 0.625)),
                                  ((7, 6, 5), (1, 1.3, 0.7))])
@pytest.mark.parametrize("split", [6, 5])
def test_reference_split_matches_reference_build(ref, n, ub, split):
    """bit-identical to the reference's own mesh code (oracle/_ref, built from its sources): the
    same points, tets with the same vertex orders, faces with the same numbering, vertex orders
    and cells, on cubes, slabs and uneven boxes (square and non-square faces: the diagonal
    choice of a square face is a floating-point tie)"""
    r = ref.box_mesh(n, (0, 0, 0), ub, split)
    m = _export4(F.Mesh.box(n, ub=ub, split=10 + split))
    for k in ("points", "cells", "face_nodes", "face_cells"):
        assert np.array_equal(m[k], r[k]), k


def test_reference_split_rejects_periodic_pairing():
    with pytest.raises(F.FvdgError):
        F.Mesh.box(4, split=16, periodic=1)


def test_library_box_is_a_valid_split_of_the_same_hexes(ref):
    """The library's 6-split and the reference's 6-split tile the same box with the same
    number of tets and faces (different diagonal choices are allowed)."""
    m = F.Mesh.box((3, 3, 2), ub=(10, 10, 2.5), split=6)
    r = ref.box_mesh((3, 3, 2), (0, 0, 0), (10, 10, 2.5), 6)
    assert m.sizes == (len(r["points"]), len(r["cells"]), len(r["face_nodes"]))
    assert np.array_equal(m.export()["points"], r["points"])
