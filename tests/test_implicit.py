"""Implicit path: the Picard-linearised block system of ImplicitConvection::assemble
(src/DG/DG_Schemes/ImplicitConvection.cpp), its GPU assembly, the block-Jacobi GMRES solve and
the backward-Euler Picard step.

Oracle pins (CPU):
- flux Jacobians: J_d(U) U = F_d(U) (the Euler flux is homogeneous of degree one) and central
  differences of F;
- interior assembly: by homogeneity, A(U) U equals the explicit Lax-Friedrichs cells + internals
  residual (itself pinned by the divergence identity) on a periodic jittered mesh;
- boundaries: wherever a type's ghost state equals the interior trace (Dirichlet and WallTN with
  the traces as data, WallTD at internal energy 300, Pseudo3DX/Y/Z with that momentum zero),
  A(U) U - rhs equals the explicit residual with Dirichlet data equal to the traces (F(U) . n on
  the boundary).
GPU parity: blocks, right-hand side and products against the oracle (1e-12 of the largest
entry), orders 0-5; the GMRES solution against a dense solve; the Picard step against the same
iteration composed from oracle assemblies and dense solves.
"""
import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX, project_pointwise, rel_err

GAMMA = 1.4


def apply_blocks(m, diag, off, U):
    """A U from the oracle's per-element diagonal and per-face off-diagonal blocks"""
    y = np.einsum("eab,eb->ea", diag, U)
    for f, (c0, c1) in enumerate(m["face_cells"]):
        if c1 >= 0:
            y[c0] += off[f, 0] @ U[c1]
            y[c1] += off[f, 1] @ U[c0]
    return y


def face_traces_at_data_faces(oracle, m, order, U, bc, kinds=(0, 9, 10)):
    """interior traces at the face points of the boundary faces whose type takes data, (faces*Qf, 5)"""
    og = oracle.geometry(m)
    fp, _ = oracle.quadrature(order, 1)
    N = F.num_basis(order)
    rows = []
    for f in range(len(m["face_nodes"])):
        if m["face_cells"][f, 1] >= 0 or bc[f] not in kinds:
            continue
        c = m["face_cells"][f, 0]
        vt = og["vert"][f, 0]
        for u, v in fp:
            lam = np.zeros(4)
            lam[vt[0]] += 1 - u - v
            lam[vt[1]] += u
            lam[vt[2]] += v
            phi, _ = oracle.basis(order, *lam[1:])
            rows.append(phi @ U[c].reshape(N, 5))
    return np.array(rows).reshape(-1, 5)


def explicit_with_traces(oracle, m, order, U, bc):
    """explicit LF residual with every boundary face Dirichlet and its data = the interior traces"""
    bnd = m["face_cells"][:, 1] < 0
    dbc = np.where(bnd, 0, bc).astype(np.uint8)
    UD = face_traces_at_data_faces(oracle, m, order, U, dbc, kinds=(0,))
    return oracle.residual(m, order, U, flux="lf", bc=dbc, UD=UD)


# ---------------------------------------------------------------- oracle pins (CPU)
def test_jacobian_is_the_flux_derivative(oracle):
    rng = np.random.default_rng(1)
    for _ in range(20):
        U = np.array([rng.uniform(0.5, 2), *rng.normal(0, 1, 3), 0])
        U[4] = rng.uniform(1, 3) / (GAMMA - 1) + 0.5 * (U[1:4] ** 2).sum() / U[0]
        J = oracle.jacobian(U, GAMMA)
        Fx = oracle.flux(U, GAMMA)  # (5, 3)
        assert np.allclose(np.einsum("dvu,u->vd", J, U), Fx, rtol=1e-13, atol=1e-13)
        h = 1e-6
        for u in range(5):
            dU = np.zeros(5)
            dU[u] = h * max(1.0, abs(U[u]))
            fd = (oracle.flux(U + dU, GAMMA) - oracle.flux(U - dU, GAMMA)) / (2 * dU[u])
            assert np.allclose(J[:, :, u].T, fd, rtol=1e-7, atol=1e-7), u


@pytest.mark.parametrize("order", [1, 2])
def test_interior_blocks_reproduce_the_explicit_residual(oracle, order):
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=5)
    m = mesh.export()
    U = mesh.project(order, "vortex", VORTEX)
    diag, off, rhs = oracle.implicit_assemble(m, order, U)
    assert not rhs.any()
    R = oracle.residual(m, order, U, flux="lf", parts=3)
    assert rel_err(apply_blocks(m, diag, off, U), R) < 1e-12


def bounded_case(oracle, order, kind):
    mesh = F.Mesh.box(2, ub=(10, 10, 10), periodic=0, jitter=0.1, seed=6)
    m = mesh.export()
    bnd = m["face_cells"][:, 1] < 0
    bc = np.full(len(bnd), 255, np.uint8)
    bc[bnd] = kind
    if kind == 9:  # WallTD: ghost = interior where rho e_internal = 300 (a quadratic energy, exact from P2)
        def fn(x):
            mom = np.array([0.4 + 0.05 * x[0], -0.3 + 0.02 * x[1], 0.2 - 0.03 * x[2]])
            return np.array([1.2, *mom, 300.0 + 0.5 * (mom ** 2).sum() / 1.2])
        U = project_pointwise(oracle, m, order, fn)
    else:
        U = mesh.project(order, "vortex", VORTEX)
        if kind in (3, 4, 5):  # mirror of momentum kind - 2 (x, y, z): that momentum zero
            N = U.shape[1] // 5
            U.reshape(-1, N, 5)[:, :, kind - 2] = 0.0
    return mesh, m, bc, U


@pytest.mark.parametrize("kind", [0, 3, 4, 5, 9, 10])
@pytest.mark.parametrize("order", [1, 2])
def test_boundary_blocks_where_the_ghost_is_the_trace(oracle, order, kind):
    if kind == 9 and order < 2:
        pytest.skip("the WallTD state needs a quadratic energy")
    mesh, m, bc, U = bounded_case(oracle, order, kind)
    UD = face_traces_at_data_faces(oracle, m, order, U, bc)
    diag, off, rhs = oracle.implicit_assemble(m, order, U, bc=bc, UD=UD if len(UD) else None)
    lhs = apply_blocks(m, diag, off, U) - rhs
    assert rel_err(lhs, explicit_with_traces(oracle, m, order, U, bc)) < 1e-11


def test_other_boundary_types_contribute_nothing(oracle):
    mesh, m, bc, U = bounded_case(oracle, 1, 0)
    bc[bc == 0] = 8  # Wall: no term in the reference's assemble_boundarys
    diag, off, rhs = oracle.implicit_assemble(m, 1, U, bc=bc)
    assert not rhs.any()
    assert rel_err(apply_blocks(m, diag, off, U), oracle.residual(m, 1, U, flux="lf", parts=3)) < 1e-12


def test_host_errors_without_gpu_calls():
    L = F.lib()
    assert L.fvdg_implicit_assemble(None, 7) == F.FVDG_ERR_ARG
    assert L.fvdg_implicit_euler_step(None, 0.1, 10, 1e-8, 1e-10, 30, 100, None, None, None) == F.FVDG_ERR_ARG
    rp = np.array([0, 1], np.int64)
    with pytest.raises(F.FvdgError):
        F.bsr_solve(rp, np.array([0]), np.ones((1, 2, 2)), np.ones(2), restart=0)


# ---------------------------------------------------------------- GPU parity
def block_dict_device(out):
    D = {}
    for e in range(len(out["diag"])):
        D[(e, e)] = D.get((e, e), 0) + out["diag"][e]
        for l in range(4):
            c = int(out["cols"][e, l])
            if c >= 0:
                D[(e, c)] = D.get((e, c), 0) + out["off"][e, l]
    return D


def block_dict_oracle(m, diag, off):
    D = {}
    for e in range(len(diag)):
        D[(e, e)] = D.get((e, e), 0) + diag[e]
    for f, (c0, c1) in enumerate(m["face_cells"]):
        if c1 >= 0:
            D[(c0, c1)] = D.get((c0, c1), 0) + off[f, 0]
            D[(c1, c0)] = D.get((c1, c0), 0) + off[f, 1]
    return D


def mixed_bc_case(order):
    # one hexahedron (6 tets, 12 boundary faces: every boundary type) from P4 on, where the blocks
    # are 175 x 175 and 280 x 280; 2^3 hexahedra below
    mesh = F.Mesh.box(2 if order <= 3 else 1, ub=(10, 10, 10), periodic=0, jitter=0.1, seed=8)
    m = mesh.export()
    bnd = np.nonzero(m["face_cells"][:, 1] < 0)[0]
    bc = mesh.boundary_types()
    bc[bnd] = np.array([0, 3, 4, 5, 9, 10, 8], np.uint8)[np.arange(len(bnd)) % 7]
    mesh.set_boundary_types(bc)
    U = mesh.project(order, "vortex", VORTEX)
    return mesh, m, mesh.boundary_types(), U


def dense(D, E, NB):
    A = np.zeros((E * NB, E * NB))
    for (r, c), B in D.items():
        A[r * NB:(r + 1) * NB, c * NB:(c + 1) * NB] += B
    return A


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["periodic", "mixed"])
@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
def test_gpu_blocks_match_oracle(solver, oracle, order, case):
    print(f"[case] implicit blocks {case} P{order}", flush=True)
    if case == "periodic":
        n = 3 if order <= 2 else 2
        mesh = F.Mesh.box(n, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=5)
        m, bc, U = mesh.export(), None, mesh.project(order, "vortex", VORTEX)
    else:
        mesh, m, bc, U = mixed_bc_case(order)
    rng = np.random.default_rng(order)
    U = U * (1 + 0.05 * rng.standard_normal(U.shape))
    solver.set_scheme(order, flux="lf").set_mesh(mesh)
    solver.upload(U)
    UD = None
    ids, xyz = solver.dirichlet_points()
    if len(ids):
        UD = np.array([oracle.vortex_state(VORTEX, 0.0, x) for x in xyz.reshape(-1, 3)])
        solver.set_dirichlet_states(UD.reshape(len(ids), -1, 5))
    solver.implicit_assemble()
    out = solver.implicit_blocks()
    diag, off, rhs = oracle.implicit_assemble(m, order, U, bc=bc, UD=UD)
    Dd, Do = block_dict_device(out), block_dict_oracle(m, diag, off)
    assert Dd.keys() == Do.keys()
    scale = max(np.abs(B).max() for B in Do.values())
    assert max(np.abs(Dd[k] - Do[k]).max() for k in Do) < 1e-12 * scale
    assert np.abs(out["rhs"] - rhs).max() <= 1e-12 * max(np.abs(rhs).max(), 1e-300)
    # products through the device matvec, with and without the mass shift
    x = rng.standard_normal(U.shape)
    M = oracle.mass(m, order)
    ref = apply_blocks(m, diag, off, x)
    assert rel_err(solver.implicit_matvec(x), ref) < 1e-12
    assert rel_err(solver.implicit_matvec(x, shift=7.0), ref + 7.0 * M * x) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2])
def test_gpu_solve_matches_dense(solver, oracle, order):
    print(f"[case] implicit solve P{order}", flush=True)
    mesh, m, bc, U = mixed_bc_case(order)
    solver.set_scheme(order, flux="lf").set_mesh(mesh)
    solver.upload(U)
    ids, xyz = solver.dirichlet_points()
    UD = np.array([oracle.vortex_state(VORTEX, 0.0, x) for x in xyz.reshape(-1, 3)])
    solver.set_dirichlet_states(UD.reshape(len(ids), -1, 5))
    solver.implicit_assemble()
    diag, off, rhs = oracle.implicit_assemble(m, order, U, bc=bc, UD=UD)
    E, NB = U.shape
    shift = 20.0
    A = dense(block_dict_oracle(m, diag, off), E, NB) + np.diag(shift * oracle.mass(m, order).ravel())
    b = np.random.default_rng(2).standard_normal(E * NB)
    x, it, rr = solver.implicit_solve(b.reshape(E, NB), shift=shift, tol=1e-13, restart=40, max_iters=2000)
    assert rr < 1e-13 and it > 0
    xs = np.linalg.solve(A, b)
    assert rel_err(x.ravel(), xs) < 1e-9
    # the assembled right-hand side as b, from a nonzero guess
    x2, _, rr2 = solver.implicit_solve(None, x0=xs.reshape(E, NB), shift=shift, tol=1e-13, restart=40)
    assert rr2 < 1e-13 and rel_err(x2.ravel(), np.linalg.solve(A, rhs.ravel())) < 1e-9


@pytest.mark.gpu
def test_gpu_bsr_solve_matches_dense():
    print("[case] block-sparse solve", flush=True)
    rng = np.random.default_rng(3)
    nr, nb = 40, 7
    rows, cols, blocks = [0], [], []
    for r in range(nr):
        cs = sorted({r, *rng.choice(nr, 3, replace=False).tolist()})
        for c in cs:
            B = rng.standard_normal((nb, nb)) * 0.1
            if c == r:
                B += np.eye(nb) * 3
            cols.append(c)
            blocks.append(B)
        rows.append(len(cols))
    A = np.zeros((nr * nb, nr * nb))
    for r in range(nr):
        for k in range(rows[r], rows[r + 1]):
            A[r * nb:(r + 1) * nb, cols[k] * nb:(cols[k] + 1) * nb] += blocks[k]
    b = rng.standard_normal(nr * nb)
    x, it, rr = F.bsr_solve(rows, cols, np.array(blocks), b, tol=1e-14, restart=20)
    assert rr < 1e-14 and rel_err(x, np.linalg.solve(A, b)) < 1e-11
    # a system GMRES(3) cannot finish in 2 iterations reports its residual honestly
    x, it, rr = F.bsr_solve(rows, cols, np.array(blocks), b, tol=1e-14, restart=3, max_iters=2)
    assert it == 2 and rr > 1e-14 and abs(rr - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) < 1e-12


@pytest.mark.gpu
def test_gpu_backward_euler_picard_step_matches_dense_iteration(solver, oracle):
    """the Picard loop of the reference's implicit drivers, composed from oracle assemblies and
    dense solves, against fvdg_implicit_euler_step"""
    print("[case] implicit Euler step", flush=True)
    order, dt = 1, 0.05
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=5)
    m = mesh.export()
    U0 = mesh.project(order, "vortex", VORTEX)
    E, NB = U0.shape
    M = oracle.mass(m, order)
    dx, it, init = U0.copy(), 0, None
    prev = None
    for it in range(1, 30):
        diag, off, rhs = oracle.implicit_assemble(m, order, dx)
        A = dense(block_dict_oracle(m, diag, off), E, NB) + np.diag(M.ravel() / dt)
        x = np.linalg.solve(A, (rhs + M / dt * U0).ravel()).reshape(E, NB)
        delta = np.linalg.norm(x - dx)
        dx = x
        if init is None:
            init = prev = delta
        rate, rel = delta / init, delta / prev
        prev = delta
        if delta < 1e-8 or rel < 1e-8 or (it > 11 and rate > 0.95):
            break
    solver.set_scheme(order, flux="lf").set_mesh(mesh)
    solver.upload(U0)
    st = solver.implicit_euler_step(dt, max_picard=30, picard_tol=1e-8, tol=1e-13, restart=40, max_iters=4000)
    assert st["picard_iters"] == it and st["linear_iters"] > 0
    assert rel_err(solver.download(), dx) < 1e-9
    # the step conserves the cell-integral sums on the periodic mesh
    s0 = (M * U0)[:, :5].sum(0)
    s1 = (M * solver.download())[:, :5].sum(0)
    assert np.abs(s1 - s0).max() < 1e-10 * np.abs(s0).max()


@pytest.mark.gpu
@pytest.mark.parametrize("order", [0, 1, 2, 3, 4])
def test_gpu_block_jacobi_is_the_exact_inverse_of_a_block_diagonal_system(solver, oracle, order):
    """Volume blocks only (a block-diagonal system): the block-Jacobi preconditioner (register
    Gauss-Jordan for 5/20/50 unknowns per block, the LDS kernel for 100, LU factors beyond) is
    the exact inverse, so GMRES stops after one iteration on the dense per-block solution."""
    print(f"[case] block-Jacobi exact P{order}", flush=True)
    mesh = F.Mesh.box(2, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=9)
    U = mesh.project(order, "vortex", VORTEX)
    solver.set_scheme(order, flux="lf").set_mesh(mesh)
    solver.upload(U)
    solver.implicit_assemble(parts=F.PART_CELLS)
    blk = solver.implicit_blocks()
    E, NB = U.shape
    assert np.abs(blk["off"]).max() == 0.0
    shift = 3.0
    M = oracle.mass(mesh.export(), order)
    b = np.random.default_rng(order).standard_normal((E, NB))
    x, it, rr = solver.implicit_solve(b, shift=shift, tol=1e-12, restart=10, max_iters=10)
    xs = np.array([np.linalg.solve(blk["diag"][e] + shift * np.diag(M[e]), b[e]) for e in range(E)])
    assert it == 1 and rr < 1e-12
    assert rel_err(x, xs) < 1e-11
