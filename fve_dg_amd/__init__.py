"""Python binding of libfvdg.so, the MI355X explicit DG Euler solver.

Everything goes through the C ABI declared in ``include/fvdg.h`` (ctypes, numpy
arrays for the buffers).  The library must have been built
(``python -m fve_dg_amd.build``); there is no fallback: a missing library or a
failing call raises :class:`FvdgError`.

State arrays use the reference's LongVector<5N> layout: shape (E, 5N), entry
``[e, 5k+v]`` for mode k and variable v in (rho, rho u, rho v, rho w, rho E).
"""
import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libfvdg.so"
_lib = None

FLUX = {"lf": 0, "roe": 1}
PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2}
FIELD = {"vortex": 0, "tgv": 1, "const": 2}
PART_CELLS, PART_INTERNALS, PART_BOUNDARYS, PART_ALL = 1, 2, 4, 7
BC_DIRICHLET, BC_PSEUDO3DZ = 0, 5
LIMIT_BOUNDS, LIMIT_POSITIVITY = 1, 2  # PositiveLimiter apply_1, apply_2
FVDG_BC_DIRICHLET, FVDG_BC_PSEUDO3DZ = BC_DIRICHLET, BC_PSEUDO3DZ
# status codes of include/fvdg.h
FVDG_OK, FVDG_ERR_ARG, FVDG_ERR_HIP, FVDG_ERR_MESH = 0, -1, -2, -3
FVDG_ERR_STATE, FVDG_ERR_UNSUPPORTED, FVDG_ERR_NONFINITE = -4, -5, -6

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_u8 = C.POINTER(C.c_uint8)
_vp = C.c_void_p


class FvdgError(RuntimeError):
    """A libfvdg call returned a nonzero status."""

    def __init__(self, status, msg):
        super().__init__(f"libfvdg status {status}: {msg}")
        self.status = status


def _sig(lib, name, args, res=C.c_int):
    fn = getattr(lib, name)
    fn.argtypes = args
    fn.restype = res


def lib():
    """Load libfvdg.so (once)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise FvdgError(-1, f"{_LIB_PATH} is not built; run `python -m fve_dg_amd.build`")
        L = C.CDLL(str(_LIB_PATH))
        _sig(L, "fvdg_version", [], C.c_char_p)
        _sig(L, "fvdg_last_error", [], C.c_char_p)
        _sig(L, "fvdg_num_basis", [C.c_int])
        _sig(L, "fvdg_basis_eval", [C.c_int, C.c_double, C.c_double, C.c_double, _d, _d])
        _sig(L, "fvdg_quadrature", [C.c_int, C.POINTER(C.c_int), _d, _d, C.POINTER(C.c_int), _d, _d])
        _sig(L, "fvdg_mesh_create_box", [C.POINTER(_vp), _d, _d, _i64, C.c_int, C.c_int, C.c_double, C.c_uint64])
        _sig(L, "fvdg_mesh_create", [C.POINTER(_vp), C.c_int64, _d, C.c_int64, _i64, C.c_int64, _i64, _i64, _i64])
        _sig(L, "fvdg_mesh_sizes", [_vp, _i64, _i64, _i64])
        _sig(L, "fvdg_mesh_export", [_vp, _d, _i64, _i64, _i64, _i64])
        _sig(L, "fvdg_mesh_geometry", [_vp, _d, _d, _d, _d, _d])
        _sig(L, "fvdg_mesh_set_boundary_types", [_vp, _u8])
        _sig(L, "fvdg_mesh_get_boundary_types", [_vp, _u8])
        _sig(L, "fvdg_mesh_tag_boundaries_pseudo3dz", [_vp])
        _sig(L, "fvdg_mesh_destroy", [_vp], None)
        _sig(L, "fvdg_mesh_read_fluent", [C.POINTER(_vp), C.c_char_p])
        _sig(L, "fvdg_mesh_face_zones", [_vp, C.POINTER(C.c_int32)])
        _sig(L, "fvdg_mesh_zones", [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p,
                                    C.c_int])
        _sig(L, "fvdg_mesh_set_zone_boundary_type", [_vp, C.c_int32, C.c_uint8])
        _sig(L, "fvdg_solution_write", [_vp, C.c_int, _d, C.c_int, _d, C.c_int, C.c_double, C.c_char_p])
        _sig(L, "fvdg_solution_read", [_vp, C.c_int, C.c_char_p, _d])
        _sig(L, "fvdg_checkpoint_write", [C.c_char_p, C.c_int, C.c_int64, _d, C.c_double, C.c_int64])
        _sig(L, "fvdg_checkpoint_read", [C.c_char_p, C.POINTER(C.c_int), _i64, _d, _d, _i64])
        _sig(L, "fvdg_project_analytic", [_vp, C.c_int, C.c_int, _d, C.c_int, C.c_double, _d])
        _sig(L, "fvdg_analytic_error", [_vp, C.c_int, C.c_int, _d, C.c_int, C.c_double, _d, _d])
        _sig(L, "fvdg_device_count", [C.POINTER(C.c_int)])
        _sig(L, "fvdg_create", [C.POINTER(_vp), C.c_int])
        _sig(L, "fvdg_destroy", [_vp], None)
        _sig(L, "fvdg_set_scheme", [_vp, C.c_int, C.c_int, C.c_double, C.c_int])
        _sig(L, "fvdg_set_mesh", [_vp, _vp])
        _sig(L, "fvdg_state_upload", [_vp, _d])
        _sig(L, "fvdg_state_download", [_vp, _d])
        _sig(L, "fvdg_dirichlet_points", [_vp, _i64, _i64, _d])
        _sig(L, "fvdg_set_dirichlet_states", [_vp, _d])
        _sig(L, "fvdg_residual", [_vp, C.c_double, C.c_int, _d])
        _sig(L, "fvdg_mass_diagonal", [_vp, _d])
        _sig(L, "fvdg_ssprk3_step", [_vp, C.c_double, C.c_int])
        _sig(L, "fvdg_ssprk3_stage", [_vp, C.c_double, C.c_int])
        _sig(L, "fvdg_stable_dt", [_vp, C.c_double, _d])
        _sig(L, "fvdg_max_wave_speed", [_vp, _d])
        _sig(L, "fvdg_conservation_sums", [_vp, _d])
        _sig(L, "fvdg_time_stages", [_vp, C.c_double, C.c_int, _d])
        _sig(L, "fvdg_stage_kernels", [_vp, C.c_char_p, C.c_int])
        _sig(L, "fvdg_flow_diagnostics", [_vp, _d])
        _ip = C.POINTER(C.c_int)
        _sig(L, "fvdg_implicit_assemble", [_vp, C.c_int])
        _sig(L, "fvdg_implicit_blocks", [_vp, _d, _d, C.POINTER(C.c_int32), _d])
        _sig(L, "fvdg_implicit_matvec", [_vp, C.c_double, _d, _d])
        _sig(L, "fvdg_implicit_solve", [_vp, C.c_double, _d, _d, C.c_double, C.c_int, C.c_int, _ip, _d])
        _sig(L, "fvdg_implicit_euler_step", [_vp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _ip,
                                             _ip, _d])
        _sig(L, "fvdg_bsr_solve", [C.c_int, C.c_int64, C.c_int, _i64, _i64, _d, _d, _d, C.c_double, C.c_int, C.c_int,
                                   _ip, _d])
        _sig(L, "fvdg_limiter_bounds", [_vp, _d, C.c_int])
        _sig(L, "fvdg_limiter_apply", [_vp, _d, C.c_int])
        _sig(L, "fvdg_set_stage_limiter", [_vp, C.c_int, C.c_int])
        _sig(L, "fvdg_ensemble_create", [C.POINTER(_vp), C.c_int, C.c_int, C.POINTER(_vp), C.POINTER(C.c_int),
                                         C.c_int, C.c_double, C.c_int])
        _sig(L, "fvdg_ensemble_destroy", [_vp], None)
        _sig(L, "fvdg_ensemble_num_cells", [_vp, C.c_int, _i64])
        _sig(L, "fvdg_ensemble_num_groups", [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)])
        _sig(L, "fvdg_ensemble_upload", [_vp, C.c_int, _d])
        _sig(L, "fvdg_ensemble_download", [_vp, C.c_int, _d])
        _sig(L, "fvdg_ensemble_max_wave_speeds", [_vp, _d])
        _sig(L, "fvdg_ensemble_set_dt", [_vp, _d])
        _sig(L, "fvdg_ensemble_step", [_vp, C.c_int, C.c_int])
        _sig(L, "fvdg_ensemble_time_steps", [_vp, C.c_int, C.c_int, _d])
        _sig(L, "fvdg_ensemble_autotune", [_vp, C.c_int, C.POINTER(C.c_int)])
        _sig(L, "fvdg_ensemble_set_serial", [_vp, C.c_int])
        _lib = L
    return _lib


def gpu_available():
    """True when the HIP runtime sees a device (fvdg_device_count of the C ABI)."""
    n = C.c_int(0)
    return lib().fvdg_device_count(C.byref(n)) == 0 and n.value > 0


def _check(status):
    if status != 0:
        raise FvdgError(status, lib().fvdg_last_error().decode())


def _p(a, ctype=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def num_basis(order):
    return (order + 1) * (order + 2) * (order + 3) // 6


def basis(order, x, y, z):
    """(values (N,), reference gradients (N,3)) of the modal basis at (x,y,z)."""
    n = num_basis(order)
    v, g = np.zeros(n), np.zeros((n, 3))
    st = lib().fvdg_basis_eval(order, x, y, z, _p(v), _p(g))
    if st < 0:
        _check(st)
    return v, g


def quadrature(order):
    """(volume points (Q,3), weights (Q,), face points (Qf,2), weights (Qf,)) of the order's rules."""
    nq, nf = C.c_int(), C.c_int()
    _check(lib().fvdg_quadrature(order, C.byref(nq), None, None, C.byref(nf), None, None))
    vp, vw = np.zeros((nq.value, 3)), np.zeros(nq.value)
    fp, fw = np.zeros((nf.value, 2)), np.zeros(nf.value)
    _check(lib().fvdg_quadrature(order, None, _p(vp), _p(vw), None, _p(fp), _p(fw)))
    return vp, vw, fp, fw


def checkpoint_write(path, order, U, t=0.0, step=0):
    """binary checkpoint of a state (E, 5N) at time t, step number `step`"""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.ndim != 2 or U.shape[1] != 5 * num_basis(order):
        raise ValueError(f"state shape {U.shape} does not match order {order}")
    _check(lib().fvdg_checkpoint_write(str(path).encode(), order, U.shape[0], _p(U), t, step))


def checkpoint_read(path):
    """(U, order, t, step) of a binary checkpoint"""
    o, n, t, st = C.c_int(), C.c_int64(), C.c_double(), C.c_int64()
    p = str(path).encode()
    _check(lib().fvdg_checkpoint_read(p, C.byref(o), C.byref(n), None, None, None))
    U = np.zeros((n.value, 5 * num_basis(o.value)))
    _check(lib().fvdg_checkpoint_read(p, None, None, _p(U), C.byref(t), C.byref(st)))
    return U, o.value, t.value, st.value


class Mesh:
    """Host tetrahedral mesh (ComputingMesh) owned by libfvdg."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def box(cls, n, lb=(0.0, 0.0, 0.0), ub=(1.0, 1.0, 1.0), split=6, periodic=0, jitter=0.0, seed=0):
        """n[0] x n[1] x n[2] hexahedra on [lb, ub] split into tets: split 6 / 5 = this library's
        translation-invariant patterns (periodic pairing: bit d of `periodic`), split 16 / 15 =
        the reference's OrthHexMesh + split_hex6_scan / split_hex5_scan + rebuild_cell_topology
        (the same tets, vertex orders and face numbering; not periodic)"""
        n = (n, n, n) if np.isscalar(n) else tuple(n)
        h = _vp()
        lb_, ub_ = np.asarray(lb, dtype=np.float64), np.asarray(ub, dtype=np.float64)
        n_ = np.asarray(n, dtype=np.int64)
        _check(lib().fvdg_mesh_create_box(C.byref(h), _p(lb_), _p(ub_), _p(n_, C.c_int64), split, periodic,
                                          jitter, seed))
        return cls(h)

    @classmethod
    def from_arrays(cls, points, cells, face_nodes, face_cells, face_nodes_r=None):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        cl = np.ascontiguousarray(cells, dtype=np.int64)
        fn = np.ascontiguousarray(face_nodes, dtype=np.int64)
        fc = np.ascontiguousarray(face_cells, dtype=np.int64)
        fr = None if face_nodes_r is None else np.ascontiguousarray(face_nodes_r, dtype=np.int64)
        h = _vp()
        _check(lib().fvdg_mesh_create(C.byref(h), len(pts), _p(pts), len(cl), _p(cl, C.c_int64), len(fn),
                                      _p(fn, C.c_int64), _p(fc, C.c_int64),
                                      None if fr is None else _p(fr, C.c_int64)))
        return cls(h)

    @classmethod
    def read_fluent(cls, path):
        """Fluent ASCII .msh file -> conforming tetrahedral mesh (non-tet cells split around
        their centroids); boundary faces typed from their zones."""
        h = _vp()
        _check(lib().fvdg_mesh_read_fluent(C.byref(h), str(path).encode()))
        return cls(h)

    def zones(self):
        """{zone id: (Fluent boundary-condition code, name)} of a mesh read from a file"""
        n = C.c_int()
        _check(lib().fvdg_mesh_zones(self._h, C.byref(n), None, None, None, 0))
        ids = (C.c_int32 * n.value)()
        bct = (C.c_int32 * n.value)()
        names = C.create_string_buffer(64 * max(n.value, 1))
        _check(lib().fvdg_mesh_zones(self._h, C.byref(n), ids, bct, names, 64))
        return {ids[i]: (bct[i], names.raw[64 * i: 64 * (i + 1)].split(b"\0")[0].decode()) for i in range(n.value)}

    def face_zones(self):
        z = np.zeros(self.sizes[2], dtype=np.int32)
        _check(lib().fvdg_mesh_face_zones(self._h, z.ctypes.data_as(C.POINTER(C.c_int32))))
        return z

    def set_zone_boundary_type(self, zone, bc):
        _check(lib().fvdg_mesh_set_zone_boundary_type(self._h, int(zone), int(bc)))

    def write_solution(self, path, order, U, field=None, params=(), t=0.0):
        """the reference's T_k_N_n.txt layout: x y z (U_h U_s) x 5 per cell and volume point"""
        U = np.ascontiguousarray(U, dtype=np.float64)
        prm = np.ascontiguousarray(params, dtype=np.float64)
        _check(lib().fvdg_solution_write(self._h, order, _p(U), -1 if field is None else FIELD[field],
                                         _p(prm) if prm.size else None, prm.size, t, str(path).encode()))

    def read_solution(self, path, order):
        U = np.zeros((self.n_cells, 5 * num_basis(order)))
        _check(lib().fvdg_solution_read(self._h, order, str(path).encode(), _p(U)))
        return U

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fvdg_mesh_destroy(self._h)
            self._h = None

    @property
    def sizes(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().fvdg_mesh_sizes(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    @property
    def n_cells(self):
        return self.sizes[1]

    def export(self):
        npnt, nc, nf = self.sizes
        out = dict(points=np.zeros((npnt, 3)), cells=np.zeros((nc, 4), np.int64),
                   face_nodes=np.zeros((nf, 3), np.int64), face_cells=np.zeros((nf, 2), np.int64),
                   face_nodes_r=np.zeros((nf, 3), np.int64))
        _check(lib().fvdg_mesh_export(self._h, _p(out["points"]), _p(out["cells"], C.c_int64),
                                      _p(out["face_nodes"], C.c_int64), _p(out["face_cells"], C.c_int64),
                                      _p(out["face_nodes_r"], C.c_int64)))
        return out

    def geometry(self):
        _, nc, nf = self.sizes
        g = dict(inv_jac=np.zeros((nc, 3, 3)), det=np.zeros(nc), h=np.zeros(nc),
                 normal=np.zeros((nf, 3)), area=np.zeros(nf))
        _check(lib().fvdg_mesh_geometry(self._h, _p(g["inv_jac"]), _p(g["det"]), _p(g["h"]), _p(g["normal"]),
                                        _p(g["area"])))
        return g

    def boundary_types(self):
        bc = np.zeros(self.sizes[2], np.uint8)
        _check(lib().fvdg_mesh_get_boundary_types(self._h, _p(bc, C.c_uint8)))
        return bc

    def set_boundary_types(self, bc):
        bc = np.ascontiguousarray(bc, dtype=np.uint8)
        _check(lib().fvdg_mesh_set_boundary_types(self._h, _p(bc, C.c_uint8)))

    def tag_pseudo3dz(self):
        _check(lib().fvdg_mesh_tag_boundaries_pseudo3dz(self._h))

    def project(self, order, field, params, t=0.0):
        prm = np.ascontiguousarray(params, dtype=np.float64)
        U = np.zeros((self.n_cells, 5 * num_basis(order)))
        _check(lib().fvdg_project_analytic(self._h, order, FIELD[field], _p(prm), len(prm), t, _p(U)))
        return U

    def analytic_error(self, order, field, params, t, U):
        prm = np.ascontiguousarray(params, dtype=np.float64)
        U = np.ascontiguousarray(U, dtype=np.float64)
        err = np.zeros(5)
        _check(lib().fvdg_analytic_error(self._h, order, FIELD[field], _p(prm), len(prm), t, _p(U), _p(err)))
        return err


class Solver:
    """One GPU solver context: scheme + mesh + device state."""

    def __init__(self, device=0):
        h = _vp()
        _check(lib().fvdg_create(C.byref(h), device))
        self._h = h
        self.order = None
        self.n_cells = None

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fvdg_destroy(self._h)
            self._h = None

    __del__ = close

    def set_scheme(self, order, flux="lf", gamma=1.4, precision="fp64"):
        _check(lib().fvdg_set_scheme(self._h, order, FLUX[flux], gamma, PRECISION[precision]))
        self.order = order
        return self

    def set_mesh(self, mesh):
        _check(lib().fvdg_set_mesh(self._h, mesh._h))
        self.n_cells = mesh.n_cells
        return self

    @property
    def nv(self):
        return 5 * num_basis(self.order)

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (self.n_cells, self.nv):
            raise ValueError(f"state shape {U.shape} != {(self.n_cells, self.nv)}")
        _check(lib().fvdg_state_upload(self._h, _p(U)))

    def download(self):
        U = np.zeros((self.n_cells, self.nv))
        _check(lib().fvdg_state_download(self._h, _p(U)))
        return U

    def residual(self, t=0.0, parts=PART_ALL):
        R = np.zeros((self.n_cells, self.nv))
        _check(lib().fvdg_residual(self._h, t, parts, _p(R)))
        return R

    def mass(self):
        M = np.zeros((self.n_cells, self.nv))
        _check(lib().fvdg_mass_diagonal(self._h, _p(M)))
        return M

    def dirichlet_points(self):
        n = C.c_int64()
        _check(lib().fvdg_dirichlet_points(self._h, C.byref(n), None, None))
        qf = len(quadrature(self.order)[3])
        ids, xyz = np.zeros(n.value, np.int64), np.zeros((n.value, qf, 3))
        _check(lib().fvdg_dirichlet_points(self._h, C.byref(n), _p(ids, C.c_int64), _p(xyz)))
        return ids, xyz

    def set_dirichlet_states(self, UD):
        UD = np.ascontiguousarray(UD, dtype=np.float64)
        _check(lib().fvdg_set_dirichlet_states(self._h, _p(UD)))

    def ssprk3(self, dt, n_steps=1):
        _check(lib().fvdg_ssprk3_step(self._h, dt, n_steps))

    def ssprk3_stage(self, dt, stage):
        """stage 0, 1 or 2 of one SSP-RK3 step (Dirichlet states may be reset in between)"""
        _check(lib().fvdg_ssprk3_stage(self._h, dt, stage))

    def stable_dt(self, cfl=0.3):
        """cfl * h_min / ((2p+1) lambda_max) of the device state"""
        out = C.c_double()
        _check(lib().fvdg_stable_dt(self._h, cfl, C.byref(out)))
        return out.value

    def max_wave_speed(self):
        out = C.c_double()
        _check(lib().fvdg_max_wave_speed(self._h, C.byref(out)))
        return out.value

    def conservation_sums(self):
        out = np.zeros(5)
        _check(lib().fvdg_conservation_sums(self._h, _p(out)))
        return out

    def stage_kernels(self):
        """dict slot -> kernel name ('volume', 'flux', 'test'; '' = nothing launched) of one stage"""
        buf = C.create_string_buffer(512)
        _check(lib().fvdg_stage_kernels(self._h, buf, 512))
        return dict(item.split("=", 1) for item in buf.value.decode().split(";"))

    def time_stages(self, dt, n_stages):
        ms = np.zeros(4)
        _check(lib().fvdg_time_stages(self._h, dt, n_stages, _p(ms)))
        return ms

    # ---- implicit path (ImplicitConvection::assemble, GMRES, backward-Euler Picard step)
    def implicit_assemble(self, parts=PART_ALL):
        """Picard blocks linearised at the device state (kept on the device)"""
        _check(lib().fvdg_implicit_assemble(self._h, parts))
        return self

    def implicit_blocks(self):
        """host copies: diag (E, NB, NB), off (E, 4, NB, NB), cols (E, 4), rhs (E, NB)"""
        E, NB = self.n_cells, self.nv
        out = dict(diag=np.zeros((E, NB, NB)), off=np.zeros((E, 4, NB, NB)), cols=np.zeros((E, 4), np.int32),
                   rhs=np.zeros((E, NB)))
        _check(lib().fvdg_implicit_blocks(self._h, _p(out["diag"]), _p(out["off"]), _p(out["cols"], C.c_int32),
                                          _p(out["rhs"])))
        return out

    def implicit_matvec(self, x, shift=0.0):
        """(A + shift M) x"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(lib().fvdg_implicit_matvec(self._h, shift, _p(x), _p(y)))
        return y

    def implicit_solve(self, b=None, x0=None, shift=0.0, tol=1e-12, restart=30, max_iters=1000):
        """solve (A + shift M) x = b (None: the assembled rhs); returns (x, iterations, relative residual)"""
        x = np.zeros((self.n_cells, self.nv)) if x0 is None else np.array(x0, dtype=np.float64, order="C")
        b_ = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        it, rr = C.c_int(), C.c_double()
        _check(lib().fvdg_implicit_solve(self._h, shift, _p(b_), _p(x), tol, restart, max_iters, C.byref(it),
                                         C.byref(rr)))
        return x, it.value, rr.value

    def implicit_euler_step(self, dt, max_picard=100, picard_tol=1e-8, tol=1e-10, restart=30, max_iters=1000):
        """one backward-Euler Picard step of the device state; returns the iteration counts"""
        pi, li, d = C.c_int(), C.c_int(), C.c_double()
        _check(lib().fvdg_implicit_euler_step(self._h, dt, max_picard, picard_tol, tol, restart, max_iters,
                                              C.byref(pi), C.byref(li), C.byref(d)))
        return dict(picard_iters=pi.value, linear_iters=li.value, delta=d.value)

    DIAGNOSTICS = ("volume", "mass", "kinetic_energy", "enstrophy", "total_energy", "min_rho", "min_p", "max_mach")

    def flow_diagnostics(self):
        """dict of the device state's integrals (volume, mass, kinetic energy, enstrophy, total
        energy) and point extremes (min rho, min p, max Mach)"""
        out = np.zeros(8)
        _check(lib().fvdg_flow_diagnostics(self._h, _p(out)))
        return dict(zip(self.DIAGNOSTICS, out.tolist()))

    # ---- limiter (PositiveLimiter.h: constructMinMax, apply / apply_1 / apply_2)
    def limiter_bounds(self, U_prev=None, neighbour_averages=False):
        """bounds from U_prev (host array) or, with None, from the device state"""
        Up = None
        if U_prev is not None:
            Up = np.ascontiguousarray(U_prev, dtype=np.float64)
            if Up.shape != (self.n_cells, self.nv):
                raise ValueError(f"state shape {Up.shape} != {(self.n_cells, self.nv)}")
        _check(lib().fvdg_limiter_bounds(self._h, _p(Up), int(bool(neighbour_averages))))

    def limit(self, U=None, parts=LIMIT_BOUNDS | LIMIT_POSITIVITY):
        """limited copy of U (host array), or, with None, the device state limited in place"""
        if U is None:
            _check(lib().fvdg_limiter_apply(self._h, None, parts))
            return None
        out = np.array(U, dtype=np.float64, copy=True, order="C")
        if out.shape != (self.n_cells, self.nv):
            raise ValueError(f"state shape {out.shape} != {(self.n_cells, self.nv)}")
        _check(lib().fvdg_limiter_apply(self._h, _p(out), parts))
        return out

    def set_stage_limiter(self, parts=LIMIT_BOUNDS | LIMIT_POSITIVITY, neighbour_averages=False):
        """limit every SSP-RK3 stage result (parts 0: off)"""
        _check(lib().fvdg_set_stage_limiter(self._h, parts, int(bool(neighbour_averages))))
        return self


def bsr_solve(row_ptr, cols, blocks, b, x0=None, tol=1e-12, restart=30, max_iters=1000, device=0):
    """GMRES solve of a block-sparse system (CSR of square blocks, row-major) on the GPU;
    returns (x, iterations, relative residual)"""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    blocks = np.ascontiguousarray(blocks, dtype=np.float64)
    nb = blocks.shape[1]
    n_rows = len(row_ptr) - 1
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(n_rows * nb)
    x = np.zeros(n_rows * nb) if x0 is None else np.array(x0, dtype=np.float64).reshape(n_rows * nb)
    it, rr = C.c_int(), C.c_double()
    _check(lib().fvdg_bsr_solve(device, n_rows, nb, _p(row_ptr, C.c_int64), _p(cols, C.c_int64), _p(blocks), _p(b),
                                _p(x), tol, restart, max_iters, C.byref(it), C.byref(rr)))
    return x, it.value, rr.value


class Ensemble:
    """Many independent problems (meshes of uneven size, orders 0..5) stepped together on
    one GPU: per order one union mesh, per problem its own time step, optional HIP graph."""

    def __init__(self, meshes, orders, flux="lf", gamma=1.4, precision="fp64", device=0):
        self.n = len(meshes)
        self.orders = [int(o) for o in orders]
        self._meshes = list(meshes)  # keep the host meshes alive for the caller's convenience
        arr = (_vp * self.n)(*[m._h for m in meshes])
        ords = (C.c_int * self.n)(*self.orders)
        h = _vp()
        _check(lib().fvdg_ensemble_create(C.byref(h), device, self.n, arr, ords, FLUX[flux], gamma,
                                          PRECISION[precision]))
        self._h = h
        self.n_cells = []
        for p in range(self.n):
            nc = C.c_int64()
            _check(lib().fvdg_ensemble_num_cells(self._h, p, C.byref(nc)))
            self.n_cells.append(nc.value)

    def groups(self):
        n = C.c_int()
        orders = (C.c_int * 8)()
        _check(lib().fvdg_ensemble_num_groups(self._h, C.byref(n), orders))
        return list(orders[: n.value])

    def upload(self, p, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (self.n_cells[p], 5 * num_basis(self.orders[p])):
            raise FvdgError(FVDG_ERR_ARG, f"state of problem {p} has shape {U.shape}")
        _check(lib().fvdg_ensemble_upload(self._h, p, _p(U)))

    def download(self, p):
        U = np.zeros((self.n_cells[p], 5 * num_basis(self.orders[p])))
        _check(lib().fvdg_ensemble_download(self._h, p, _p(U)))
        return U

    def max_wave_speeds(self):
        out = np.zeros(self.n)
        _check(lib().fvdg_ensemble_max_wave_speeds(self._h, _p(out)))
        return out

    def set_dt(self, dt):
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n,)))
        _check(lib().fvdg_ensemble_set_dt(self._h, _p(dt)))

    def step(self, n_steps=1, graph=True):
        _check(lib().fvdg_ensemble_step(self._h, n_steps, int(graph)))

    def time_steps(self, n_steps, graph=True):
        ms = C.c_double()
        _check(lib().fvdg_ensemble_time_steps(self._h, n_steps, int(graph), C.byref(ms)))
        return ms.value

    def autotune(self, trial_steps=2):
        """per order: {'elem_fused': bool, 'face_flux': 'tile' | 'dpp' | 'lds', 'flux_threads': block size}"""
        ng = len(self.groups())
        ch = (C.c_int * (2 * ng))()
        _check(lib().fvdg_ensemble_autotune(self._h, trial_steps, ch))
        kinds = {-1: "tile", 0: "lds"}
        return {o: {"elem_fused": bool(ch[2 * i]), "face_flux": kinds.get(ch[2 * i + 1], "dpp"),
                    "flux_threads": ch[2 * i + 1] if ch[2 * i + 1] > 0 else None}
                for i, o in enumerate(self.groups())}

    def set_serial(self, serial=True):
        """every order group on one stream after another (comparison) instead of concurrent groups"""
        _check(lib().fvdg_ensemble_set_serial(self._h, int(bool(serial))))

    def close(self):
        if getattr(self, "_h", None):
            lib().fvdg_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
