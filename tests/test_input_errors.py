"""Host-side input errors: empty, zero-size and inconsistent inputs come back as status codes with a
message (FvdgError), never as a device call.  CPU only."""
import numpy as np
import pytest

import fve_dg_amd as F

VORTEX = [1.4, 5.0, 5.0, 5.0, 1.0, 1.0, 10.0, 10.0]


def test_empty_mesh_arrays_are_rejected():
    with pytest.raises(F.FvdgError, match="empty"):
        F.Mesh.from_arrays(np.zeros((0, 3)), np.zeros((0, 4), np.int64), np.zeros((0, 3), np.int64),
                           np.zeros((0, 2), np.int64))


@pytest.mark.parametrize("n", [0, (2, 0, 2), (2, 2, -1)])
def test_box_needs_a_hexahedron_per_axis(n):
    with pytest.raises(F.FvdgError, match="n must be >= 1"):
        F.Mesh.box(n, ub=(1, 1, 1))


def test_box_bounds_must_be_ordered():
    with pytest.raises(F.FvdgError, match="ub must exceed lb"):
        F.Mesh.box(2, lb=(1, 1, 1), ub=(0, 0, 0))


def test_cell_faces_must_be_present():
    """one tetrahedron with one of its four faces given: the mesh is rejected"""
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    with pytest.raises(F.FvdgError, match="missing face"):
        F.Mesh.from_arrays(pts, np.array([[0, 1, 2, 3]]), np.array([[0, 1, 2]]), np.array([[0, -1]]))


@pytest.mark.parametrize("order", [-1, 6])
def test_projection_order_range(order):
    m = F.Mesh.box(2, ub=(10, 10, 10), periodic=7)
    with pytest.raises(F.FvdgError, match="order must be 0..5"):
        m.project(order, "vortex", VORTEX)


def test_smallest_periodic_box_steps_on_the_host_side():
    """two hexahedra per axis (48 tets) is the smallest periodic box whose opposite faces pair
    distinct cells; its mesh and projection build on the host"""
    m = F.Mesh.box(2, ub=(10, 10, 10), periodic=7)
    assert m.n_cells == 48
    U = m.project(3, "vortex", VORTEX)
    assert U.shape == (48, 100) and np.isfinite(U).all()
