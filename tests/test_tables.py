"""Quadrature rules and modal basis: the library's tables and the oracle's
restatement, pinned against the reference's own compiled code (oracle/_ref) and
against closed forms (monomial integrals, exact-rational orthogonality)."""
from fractions import Fraction
from math import comb, factorial

import numpy as np
import pytest

import fve_dg_amd as F
from conftest import rel_err

ORDERS = [0, 1, 2, 3, 4, 5]


def tri_exact(a, b):
    return factorial(a) * factorial(b) / factorial(a + b + 2)


def tet_exact(a, b, c):
    return factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)


@pytest.mark.parametrize("order", ORDERS)
def test_quadrature_bit_identical_to_reference(ref, oracle, order):
    vp, vw, fp, fw = F.quadrature(order)
    rp, rw = ref.quadrature(order, 0)
    assert np.array_equal(vp, rp) and np.array_equal(vw, rw)
    rp, rw = ref.quadrature(order, 1)
    assert np.array_equal(fp, rp) and np.array_equal(fw, rw)
    op, ow = oracle.quadrature(order, 0)
    assert np.array_equal(op, vp) and np.array_equal(ow, vw)


@pytest.mark.parametrize("order", ORDERS)
def test_quadrature_weights_and_exactness(order):
    vp, vw, fp, fw = F.quadrature(order)
    assert abs(vw.sum() - 1 / 6) < 1e-15 and abs(fw.sum() - 0.5) < 1e-15
    # volume rule: exact (to round-off) for every monomial up to the rule degree 2p+1 (>= 5)
    deg = {0: 5, 1: 5, 2: 5, 3: 7, 4: 9, 5: 11}[order]
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                q = (vw * vp[:, 0] ** a * vp[:, 1] ** b * vp[:, 2] ** c).sum()
                assert abs(q - tet_exact(a, b, c)) <= 2e-15 * tet_exact(a, b, c) + 1e-17, (a, b, c)
    # face rule: the degree-4 rule (P0, P1) is exact to round-off; the rotationally symmetric
    # rules of P2..P5 carry the published data's ~1e-10 relative accuracy
    fdeg = {0: 4, 1: 4, 2: 10, 3: 10, 4: 10, 5: 13}[order]
    tol = 2e-15 if order <= 1 else (2e-9 if order <= 4 else 5e-9)
    for a in range(fdeg + 1):
        for b in range(fdeg + 1 - a):
            q = (fw * fp[:, 0] ** a * fp[:, 1] ** b).sum()
            assert abs(q - tri_exact(a, b)) <= tol * tri_exact(a, b), (a, b)


def test_face_rule_is_rotation_only_symmetric():
    """Degree10Points24 maps to itself under the cyclic vertex rotations but not under
    a reflection (which is why face traces need all 24 vertex permutations)."""
    _, _, fp, fw = F.quadrature(3)
    bary = np.column_stack([1 - fp.sum(1), fp[:, 0], fp[:, 1]])

    def same_set(B):
        return all(np.min(np.abs(bary - b).max(1)) < 1e-12 for b in B)

    assert same_set(bary[:, [1, 2, 0]]) and same_set(bary[:, [2, 0, 1]])
    assert not same_set(bary[:, [0, 2, 1]])


@pytest.mark.parametrize("order", ORDERS)
def test_basis_matches_reference(ref, oracle, order):
    rng = np.random.default_rng(order)
    pts = list(F.quadrature(order)[0]) + [rng.dirichlet([1, 1, 1, 1])[:3] for _ in range(20)]
    pts += [np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0])]
    for x in pts:
        rv, rg = ref.basis(order, *x)
        lv, lg = F.basis(order, *x)
        ov, og = oracle.basis(order, *x)
        scale = max(1.0, np.abs(rv).max())
        assert np.abs(lv - rv).max() <= 1e-13 * scale and np.abs(ov - rv).max() <= 1e-13 * scale
        gs = max(1.0, np.abs(rg).max())
        assert np.abs(lg - rg).max() <= 1e-13 * gs and np.abs(og - rg).max() <= 1e-13 * gs


def _poly_mul(p, q):
    r = {}
    for (a1, b1, c1), v1 in p.items():
        for (a2, b2, c2), v2 in q.items():
            k = (a1 + a2, b1 + b2, c1 + c2)
            r[k] = r.get(k, 0) + v1 * v2
    return {k: v for k, v in r.items() if v != 0}


def _poly_add(p, q, s=1):
    r = dict(p)
    for k, v in q.items():
        r[k] = r.get(k, 0) + s * v
    return {k: v for k, v in r.items() if v != 0}


def _poly_pow(p, m):
    r = {(0, 0, 0): Fraction(1)}
    for _ in range(m):
        r = _poly_mul(r, p)
    return r


def _jacobi_poly(n, alpha, z, s):
    # s^n P_n^{(alpha,0)}((2z-s)/s) = sum_m C(n+alpha, n-m) C(n, m) (z-s)^m z^(n-m)
    acc = {}
    zs = _poly_add(z, s, -1)
    for m in range(n + 1):
        term = _poly_mul(_poly_pow(zs, m), _poly_pow(z, n - m))
        c = comb(n + alpha, n - m) * comb(n, m)
        acc = _poly_add(acc, {k: v * c for k, v in term.items()})
    return acc


def exact_basis(order):
    X, Y, Z = {(1, 0, 0): Fraction(1)}, {(0, 1, 0): Fraction(1)}, {(0, 0, 1): Fraction(1)}
    one = {(0, 0, 0): Fraction(1)}
    s = _poly_add(_poly_add(one, X, -1), Y, -1)
    r = _poly_add(one, X, -1)
    out = []
    for n in range(order + 1):
        for k in range(n + 1):
            for j in range(n - k + 1):
                i = n - j - k
                out.append(_poly_mul(_poly_mul(_jacobi_poly(k, 0, Z, s), _jacobi_poly(j, 2 * k + 1, Y, r)),
                                     _jacobi_poly(i, 2 * j + 2 * k + 2, X, one)))
    return out


def _integrate(p):
    return sum(v * Fraction(factorial(a) * factorial(b) * factorial(c), factorial(a + b + c + 3))
               for (a, b, c), v in p.items())


def test_basis_is_orthogonal_in_exact_arithmetic():
    """Gram matrix of the exact-rational Dubiner polynomials is diagonal on the unit tet,
    and the library's mass diagonal (sum_q w_q phi_k^2) equals its diagonal."""
    P = 3
    polys = exact_basis(P)
    assert len(polys) == F.num_basis(P)
    assert polys[1] == {(1, 0, 0): 4, (0, 0, 0): -1}                       # 4x - 1
    assert polys[2] == {(1, 0, 0): 1, (0, 1, 0): 3, (0, 0, 0): -1}         # x + 3y - 1
    assert polys[3] == {(1, 0, 0): 1, (0, 1, 0): 1, (0, 0, 1): 2, (0, 0, 0): -1}
    G = [[_integrate(_poly_mul(a, b)) for b in polys] for a in polys]
    for i in range(len(polys)):
        for j in range(len(polys)):
            if i != j:
                assert G[i][j] == 0
    vp, vw, _, _ = F.quadrature(P)
    Vq = np.array([F.basis(P, *x)[0] for x in vp])
    mhat = (Vq ** 2 * vw[:, None]).sum(0)
    assert rel_err(mhat, np.array([float(G[k][k]) for k in range(len(polys))])) < 1e-14
    # the exact polynomials agree with the library basis at random points
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.dirichlet([1, 1, 1, 1])[:3]
        ex = np.array([float(sum(v * x[0] ** a * x[1] ** b * x[2] ** c for (a, b, c), v in p.items()))
                       for p in polys])
        assert np.abs(F.basis(P, *x)[0] - ex).max() < 1e-13


def test_basis_gradients_by_finite_differences():
    x0 = np.array([0.21, 0.17, 0.33])
    for order in (1, 3, 4):
        _, g = F.basis(order, *x0)
        h = 1e-6
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            fd = (F.basis(order, *(x0 + e))[0] - F.basis(order, *(x0 - e))[0]) / (2 * h)
            assert np.abs(fd - g[:, d]).max() < 1e-6 * max(1, np.abs(g).max())
