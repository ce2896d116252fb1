"""Solution files in the reference's text layout (examples/Vortex/VortexBatch3.cpp:402-415) and
the restart read-back (examples/Vortex/tools.h:72-129)."""
import numpy as np

import fve_dg_amd as F
from conftest import VORTEX, rel_err


def test_write_layout_and_values(tmp_path, oracle):
    mesh = F.Mesh.box(2, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=3)
    order = 2
    U = mesh.project(order, "vortex", VORTEX)
    path = tmp_path / "T_1_N_2.txt"
    mesh.write_solution(path, order, U, field="vortex", params=VORTEX, t=0.25)
    lines = path.read_text().splitlines()
    vp, vw, _, _ = F.quadrature(order)
    assert len(lines) == mesh.n_cells * len(vw)
    rows = np.array([[float(x) for x in l.split("  ")] for l in lines])
    assert rows.shape[1] == 13
    # row (cell 1, point 2): position, DG value and exact value
    m = mesh.export()
    X = m["points"][m["cells"][1]]
    xi = vp[2]
    pos = X[0] * (1 - xi.sum()) + X[1] * xi[0] + X[2] * xi[1] + X[3] * xi[2]
    r = rows[1 * len(vw) + 2]
    assert np.allclose(r[:3], pos, rtol=1e-15, atol=1e-14)
    phi, _ = F.basis(order, *xi)
    uh = phi @ U[1].reshape(-1, 5)
    us = oracle.vortex_state(VORTEX, 0.25, pos)
    assert np.allclose(r[3::2], uh, rtol=1e-15, atol=1e-15)
    assert np.allclose(r[4::2], us, rtol=1e-15, atol=1e-15)
    # 16 significant digits, two-space separated, as the reference writes them
    def significant(tok):
        return len(tok.lstrip("-").split("e")[0].replace(".", "").lstrip("0"))

    assert all(significant(t) <= 16 for t in lines[0].split("  "))


def test_read_back_recovers_the_coefficients(tmp_path):
    mesh = F.Mesh.box(2, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=4)
    for order in (0, 1, 3, 4):
        U = mesh.project(order, "vortex", VORTEX)
        path = tmp_path / f"T_0_N_2_P{order}.txt"
        mesh.write_solution(path, order, U)
        V = mesh.read_solution(path, order)
        # the volume rule integrates phi_j phi_k exactly, so the projection of the P-space
        # point values is the identity up to the 16-digit text rounding
        assert rel_err(V, U) < 1e-14, order


def test_missing_or_short_file_is_an_error(tmp_path):
    import pytest

    mesh = F.Mesh.box(2, periodic=7)
    with pytest.raises(F.FvdgError):
        mesh.read_solution(tmp_path / "missing.txt", 1)
    (tmp_path / "short.txt").write_text("0 0 0  1 1  0 0  0 0  0 0  2.5 2.5\n")
    with pytest.raises(F.FvdgError):
        mesh.read_solution(tmp_path / "short.txt", 1)
