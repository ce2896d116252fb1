"""fvdg_stable_dt_noalloc (the time-step rule with its reduction scratch in a library-static device
array, no device allocation) against fvdg_stable_dt, through the C ABI: equal on two calls in a row,
the second on a smaller mesh (the scratch keeps the first call's partials), and on a state with a
NaN (FVDG_ERR_NONFINITE from both)."""
import ctypes as C

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX

pytestmark = pytest.mark.gpu


def noalloc_dt(solver, cfl):
    L = F.lib()
    L.fvdg_stable_dt_noalloc.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double)]
    L.fvdg_stable_dt_noalloc.restype = C.c_int
    out = C.c_double()
    return L.fvdg_stable_dt_noalloc(solver._h, cfl, C.byref(out)), out.value


def test_noalloc_dt_matches_allocating_entry_point():
    s = F.Solver(0)
    try:
        for n, order in ((24, 3), (4, 2)):  # the second call on a smaller mesh
            print(f"[case] stable_dt_noalloc n={n} P{order}", flush=True)
            mesh = F.Mesh.box(n, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=n)
            s.set_scheme(order, flux="roe").set_mesh(mesh)
            s.upload(mesh.project(order, "vortex", VORTEX))
            ref = s.stable_dt(0.3)
            rc, dt = noalloc_dt(s, 0.3)
            assert rc == 0, F.lib().fvdg_last_error()
            assert dt == ref, (dt, ref)  # a maximum does not depend on the reduction order
        U = s.download()
        U[3, 7] = np.nan
        s.upload(U)
        rc, _ = noalloc_dt(s, 0.3)
        assert rc == F.FVDG_ERR_NONFINITE
    finally:
        s.close()
