"""ctypes binding of the CPU oracle (liboracle.so) and of the reference build
(_ref/libref.so).  TEST INFRASTRUCTURE: imported only by tests/, by
``__graft_entry__.smoke()`` and by bench.py's cpu_baseline leg, always as the
checker, never as the thing measured or shipped.
"""
import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_u8 = C.POINTER(C.c_uint8)


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _sig(L, name, args, res=C.c_int):
    f = getattr(L, name)
    f.argtypes = args
    f.restype = res


def num_basis(order):
    return (order + 1) * (order + 2) * (order + 3) // 6


class Oracle:
    """The C restatement of ExplicitConvection (oracle/dg_oracle.c)."""

    def __init__(self):
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = C.CDLL(str(path))
        _sig(L, "or_quadrature", [C.c_int, C.c_int, _d, _d])
        _sig(L, "or_num_basis", [C.c_int])
        _sig(L, "or_basis", [C.c_int, C.c_double, C.c_double, C.c_double, _d, _d], None)
        _sig(L, "or_pressure", [_d, C.c_double], C.c_double)
        _sig(L, "or_flux", [_d, C.c_double, _d], None)
        _sig(L, "or_wave_speed", [_d, _d, C.c_double], C.c_double)
        _sig(L, "or_lf", [_d, _d, _d, C.c_double, _d], None)
        _sig(L, "or_roe", [_d, _d, _d, C.c_double, _d, C.c_int], None)
        _sig(L, "or_inverse3", [_d, _d], None)
        _sig(L, "or_geometry", [C.c_int64, _d, C.c_int64, _i64, C.c_int64, _i64, _i64, _i64, _d, _d, _d, _d,
                                C.POINTER(C.c_int)], None)
        _sig(L, "or_residual", [C.c_int, C.c_int, C.c_double, C.c_int64, _d, C.c_int64, _i64, C.c_int64, _i64, _i64,
                                _i64, _u8, _d, _d, C.c_int, _d, C.c_int])
        _sig(L, "or_mass", [C.c_int, C.c_int64, _d, C.c_int64, _i64, _d])
        _sig(L, "or_ssprk3", [C.c_int, C.c_int, C.c_double, C.c_int64, _d, C.c_int64, _i64, C.c_int64, _i64, _i64,
                              _i64, _u8, _d, C.c_double, C.c_int, C.c_int])
        _sig(L, "or_vortex_state", [_d, C.c_double, _d, _d], None)
        _sig(L, "or_project_vortex", [C.c_int, _d, C.c_double, C.c_int64, _d, C.c_int64, _i64, _d])
        _sig(L, "or_max_threads", [])
        _sig(L, "or_jacobian", [_d, C.c_double, _d], None)
        _sig(L, "or_implicit_assemble", [C.c_int, C.c_double, C.c_int64, _d, C.c_int64, _i64, C.c_int64, _i64, _i64,
                                         _i64, _u8, _d, _d, _d, _d, _d])
        _sig(L, "or_limiter_bounds", [C.c_int, C.c_int64, C.c_int64, _i64, _d, C.c_int, _d, _d])
        _sig(L, "or_limiter_apply", [C.c_int, C.c_int64, _d, _d, C.c_double, C.c_int, _d])
        self.L = L

    # -------------------------------------------------- pieces
    def quadrature(self, order, face):
        n = self.L.or_quadrature(order, face, None, None)
        pts, w = np.zeros((n, 2 if face else 3)), np.zeros(n)
        self.L.or_quadrature(order, face, _p(pts), _p(w))
        return pts, w

    def basis(self, order, x, y, z):
        n = num_basis(order)
        v, g = np.zeros(n), np.zeros((n, 3))
        self.L.or_basis(order, x, y, z, _p(v), _p(g))
        return v, g

    def flux(self, U, gamma=1.4):
        U = np.ascontiguousarray(U, dtype=np.float64)
        F = np.zeros(15)
        self.L.or_flux(_p(U), gamma, _p(F))
        return F.reshape(5, 3)

    def wave_speed(self, UL, UR, gamma=1.4):
        return self.L.or_wave_speed(_p(np.ascontiguousarray(UL, float)), _p(np.ascontiguousarray(UR, float)), gamma)

    def numflux(self, kind, UL, UR, n, gamma=1.4, signed_eigs=False):
        UL, UR, n = (np.ascontiguousarray(a, dtype=np.float64) for a in (UL, UR, n))
        out = np.zeros(5)
        if kind == "lf":
            self.L.or_lf(_p(UL), _p(UR), _p(n), gamma, _p(out))
        else:
            self.L.or_roe(_p(UL), _p(UR), _p(n), gamma, _p(out), int(signed_eigs))
        return out

    def inverse3(self, A):
        A = np.ascontiguousarray(A, dtype=np.float64)
        out = np.zeros(9)
        self.L.or_inverse3(_p(A), _p(out))
        return out.reshape(3, 3)

    def geometry(self, m):
        nc, nf = len(m["cells"]), len(m["face_nodes"])
        out = dict(A=np.zeros((nc, 3, 3)), det=np.zeros(nc), normal=np.zeros((nf, 3)), area=np.zeros(nf),
                   vert=np.zeros((nf, 2, 3), np.int32))
        a = _args(m)
        self.L.or_geometry(len(a["points"]), _p(a["points"]), nc, _p(a["cells"], C.c_int64), nf,
                           _p(a["face_nodes"], C.c_int64), _p(a["face_cells"], C.c_int64),
                           _p(a["face_nodes_r"], C.c_int64), _p(out["A"]), _p(out["det"]), _p(out["normal"]),
                           _p(out["area"]), out["vert"].ctypes.data_as(C.POINTER(C.c_int)))
        return out

    # -------------------------------------------------- residual and stepping
    def residual(self, m, order, U, flux="lf", gamma=1.4, bc=None, UD=None, parts=7, threads=1):
        a = _args(m)
        nc, nf = len(a["cells"]), len(a["face_nodes"])
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros_like(U)
        bc_ = None if bc is None else np.ascontiguousarray(bc, dtype=np.uint8)
        UD_ = None if UD is None else np.ascontiguousarray(UD, dtype=np.float64)
        rc = self.L.or_residual(order, 1 if flux == "roe" else 0, gamma, len(a["points"]), _p(a["points"]), nc,
                                _p(a["cells"], C.c_int64), nf, _p(a["face_nodes"], C.c_int64),
                                _p(a["face_cells"], C.c_int64), _p(a["face_nodes_r"], C.c_int64),
                                _p(bc_, C.c_uint8), _p(UD_), _p(U), parts, _p(R), threads)
        if rc:
            raise RuntimeError(f"or_residual failed ({rc})")
        return R

    def mass(self, m, order):
        a = _args(m)
        M = np.zeros((len(a["cells"]), 5 * num_basis(order)))
        self.L.or_mass(order, len(a["points"]), _p(a["points"]), len(a["cells"]), _p(a["cells"], C.c_int64), _p(M))
        return M

    def ssprk3(self, m, order, U, dt, n_steps, flux="lf", gamma=1.4, bc=None, threads=1):
        a = _args(m)
        U = np.array(U, dtype=np.float64, copy=True, order="C")
        bc_ = None if bc is None else np.ascontiguousarray(bc, dtype=np.uint8)
        self.L.or_ssprk3(order, 1 if flux == "roe" else 0, gamma, len(a["points"]), _p(a["points"]), len(a["cells"]),
                         _p(a["cells"], C.c_int64), len(a["face_nodes"]), _p(a["face_nodes"], C.c_int64),
                         _p(a["face_cells"], C.c_int64), _p(a["face_nodes_r"], C.c_int64), _p(bc_, C.c_uint8),
                         _p(U), dt, n_steps, threads)
        return U

    def vortex_state(self, params, t, x):
        P = np.ascontiguousarray(params, dtype=np.float64)
        out = np.zeros(5)
        self.L.or_vortex_state(_p(P), t, _p(np.ascontiguousarray(x, float)), _p(out))
        return out

    def project_vortex(self, m, order, params, t=0.0):
        a = _args(m)
        P = np.ascontiguousarray(params, dtype=np.float64)
        U = np.zeros((len(a["cells"]), 5 * num_basis(order)))
        self.L.or_project_vortex(order, _p(P), t, len(a["points"]), _p(a["points"]), len(a["cells"]),
                                 _p(a["cells"], C.c_int64), _p(U))
        return U

    def max_threads(self):
        return self.L.or_max_threads()

    # -------------------------------------------------- implicit (ImplicitConvection.cpp)
    def jacobian(self, U, gamma=1.4):
        """(3, 5, 5): dF_d/dU"""
        J = np.zeros((3, 25))
        self.L.or_jacobian(_p(np.ascontiguousarray(U, float)), gamma, _p(J))
        return J.reshape(3, 5, 5)

    def implicit_assemble(self, m, order, U, gamma=1.4, bc=None, UD=None):
        """Picard blocks at U: diag (E, NB, NB), off (F, 2, NB, NB) = blocks (L, R) and (R, L) of
        each interior face, rhs (E, NB); UD per Dirichlet / WallTD / WallTN face, (faces*Qf, 5)"""
        a = _args(m)
        U = np.ascontiguousarray(U, dtype=np.float64)
        E, NB = U.shape
        nf = len(a["face_nodes"])
        diag, off, rhs = np.zeros((E, NB, NB)), np.zeros((nf, 2, NB, NB)), np.zeros((E, NB))
        bc_ = None if bc is None else np.ascontiguousarray(bc, dtype=np.uint8)
        UD_ = None if UD is None else np.ascontiguousarray(UD, dtype=np.float64)
        rc = self.L.or_implicit_assemble(order, gamma, len(a["points"]), _p(a["points"]), E, _p(a["cells"], C.c_int64),
                                         nf, _p(a["face_nodes"], C.c_int64), _p(a["face_cells"], C.c_int64),
                                         _p(a["face_nodes_r"], C.c_int64), _p(bc_, C.c_uint8), _p(UD_), _p(U),
                                         _p(diag), _p(off), _p(rhs))
        if rc:
            raise RuntimeError(f"or_implicit_assemble failed ({rc})")
        return diag, off, rhs

    # -------------------------------------------------- limiter (PositiveLimiter.h)
    def limiter_bounds(self, m, order, U_prev, only_avg=False):
        """per-cell (max, min) bounds of constructMinMax, each (E, 5)"""
        fc = np.ascontiguousarray(m["face_cells"], dtype=np.int64)
        U = np.ascontiguousarray(U_prev, dtype=np.float64)
        E = U.shape[0]
        mx, mn = np.zeros((E, 5)), np.zeros((E, 5))
        rc = self.L.or_limiter_bounds(order, E, len(fc), _p(fc, C.c_int64), _p(U), int(only_avg), _p(mx), _p(mn))
        if rc:
            raise RuntimeError(f"or_limiter_bounds failed ({rc})")
        return mx, mn

    def limiter_apply(self, order, U, bounds, gamma=1.4, parts=3):
        """apply_1 (parts & 1) then apply_2 (parts & 2) on a copy of U"""
        out = np.array(U, dtype=np.float64, copy=True, order="C")
        mx, mn = (np.ascontiguousarray(b, dtype=np.float64) for b in bounds)
        rc = self.L.or_limiter_apply(order, out.shape[0], _p(mx), _p(mn), gamma, parts, _p(out))
        if rc:
            raise RuntimeError(f"or_limiter_apply failed ({rc})")
        return out


def _args(m):
    out = {k: np.ascontiguousarray(m[k], dtype=np.float64 if k == "points" else np.int64)
           for k in ("points", "cells", "face_nodes", "face_cells")}
    fr = m.get("face_nodes_r")
    out["face_nodes_r"] = np.ascontiguousarray(m["face_nodes"] if fr is None else fr, dtype=np.int64)
    return out


class Reference:
    """The reference's own compiled flux / basis / quadrature / mesh code (_ref/libref.so)."""

    PATH = HERE / "_ref" / "libref.so"

    @classmethod
    def available(cls):
        return cls.PATH.exists()

    def __init__(self):
        L = C.CDLL(str(self.PATH))
        _sig(L, "ref_basis", [C.c_int, C.c_double, C.c_double, C.c_double, _d, _d])
        _sig(L, "ref_quadrature", [C.c_int, C.c_int, _d, _d])
        _sig(L, "ref_flux", [C.c_int, _d, _d], None)
        _sig(L, "ref_wave_speed", [C.c_int, _d, _d], C.c_double)
        _sig(L, "ref_inverse3", [_d, _d], None)
        _sig(L, "ref_box_mesh", [_d, _d, _i64, C.c_int, _i64, _i64, _i64])
        _sig(L, "ref_box_mesh_export", [_d, _i64, _i64, _i64], None)
        self.L = L

    def basis(self, order, x, y, z):
        n = num_basis(order)
        v, g = np.zeros(n), np.zeros((n, 3))
        self.L.ref_basis(order, x, y, z, _p(v), _p(g))
        return v, g

    def quadrature(self, order, face):
        n = self.L.ref_quadrature(order, face, None, None)
        pts, w = np.zeros((n, 2 if face else 3)), np.zeros(n)
        self.L.ref_quadrature(order, face, _p(pts), _p(w))
        return pts, w

    def flux(self, U, gamma_kind=0):
        F = np.zeros(15)
        self.L.ref_flux(gamma_kind, _p(np.ascontiguousarray(U, float)), _p(F))
        return F.reshape(5, 3)

    def wave_speed(self, UL, UR, gamma_kind=0):
        return self.L.ref_wave_speed(gamma_kind, _p(np.ascontiguousarray(UL, float)), _p(np.ascontiguousarray(UR, float)))

    def inverse3(self, A):
        out = np.zeros(9)
        self.L.ref_inverse3(_p(np.ascontiguousarray(A, float)), _p(out))
        return out.reshape(3, 3)

    def box_mesh(self, n, lb=(0, 0, 0), ub=(1, 1, 1), split=6):
        lb_, ub_ = np.asarray(lb, float), np.asarray(ub, float)
        n_ = np.asarray((n, n, n) if np.isscalar(n) else n, np.int64)
        np_, nc, nf = C.c_int64(), C.c_int64(), C.c_int64()
        rc = self.L.ref_box_mesh(_p(lb_), _p(ub_), _p(n_, C.c_int64), split, C.byref(np_), C.byref(nc), C.byref(nf))
        if rc:
            raise RuntimeError(f"reference mesh build failed ({rc})")
        out = dict(points=np.zeros((np_.value, 3)), cells=np.zeros((nc.value, 4), np.int64),
                   face_nodes=np.zeros((nf.value, 3), np.int64), face_cells=np.zeros((nf.value, 2), np.int64))
        self.L.ref_box_mesh_export(_p(out["points"]), _p(out["cells"], C.c_int64), _p(out["face_nodes"], C.c_int64),
                                   _p(out["face_cells"], C.c_int64))
        return out
