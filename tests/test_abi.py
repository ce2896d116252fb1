"""The C-ABI library: it loads, exports every function include/fvdg.h declares, and its host-side
argument checks return status codes (no GPU needed; no compute call reaches a device here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import fve_dg_amd as F

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = []
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?(?:int|double|void|char|int64_t)\s*\**\s*(fvdg_\w+)\s*\(", text, re.M)
    return names


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert len(names) >= 28
    L = F.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_basis_sizes():
    L = F.lib()
    assert L.fvdg_version().decode().startswith("fvdg")
    for p in range(6):
        assert L.fvdg_num_basis(p) == (p + 1) * (p + 2) * (p + 3) // 6
    assert L.fvdg_num_basis(-1) == 0


def test_argument_errors_are_status_codes():
    L = F.lib()
    assert L.fvdg_mesh_sizes(None, None, None, None) == F.FVDG_ERR_ARG
    assert "null" in L.fvdg_last_error().decode().lower() or L.fvdg_last_error()
    assert L.fvdg_residual(None, C.c_double(0.0), 7, None) == F.FVDG_ERR_ARG
    assert L.fvdg_set_scheme(None, 3, 0, C.c_double(1.4), 0) == F.FVDG_ERR_ARG
    with pytest.raises(F.FvdgError):
        F.Mesh.box(0)
    with pytest.raises(F.FvdgError):
        F.Mesh.box(2, split=4)
    with pytest.raises(F.FvdgError):
        F.basis(9, 0.1, 0.1, 0.1)


def test_boundary_type_roundtrip():
    m = F.Mesh.box(2, periodic=0)
    bc = m.boundary_types()
    bnd = m.export()["face_cells"][:, 1] < 0
    assert (bc[bnd] == F.FVDG_BC_DIRICHLET).all()
    new = bc.copy()
    new[np.nonzero(bnd)[0][:3]] = F.FVDG_BC_PSEUDO3DZ
    m.set_boundary_types(new)
    assert np.array_equal(m.boundary_types(), new)


@pytest.mark.skipif(F.gpu_available(), reason="checks the no-device error path")
def test_context_without_a_device_fails_loudly():
    with pytest.raises(F.FvdgError) as e:
        F.Solver(0)
    assert e.value.status == F.FVDG_ERR_HIP


def test_product_path_does_not_use_the_oracle():
    """The package neither imports nor loads anything under oracle/."""
    for p in (ROOT / "fve_dg_amd").rglob("*"):
        if p.suffix in (".py", ".cpp", ".hpp", ".hip", ".h"):
            text = p.read_text()
            for pat in ("import oracle", "from oracle", "liboracle", "libref", "oracle.binding", "dg_oracle"):
                assert pat not in text, (p, pat)
