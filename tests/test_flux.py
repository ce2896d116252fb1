"""Euler flux, wave speed and numerical fluxes of the oracle: bit-identical to
the reference's compiled EulerPhysicalFlux, plus closed forms and the standard
invariants (consistency, conservation, rotational invariance, Roe's property)."""
import numpy as np
import pytest

from conftest import rel_err


def random_states(n, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.2, 3.0, n)
    u = rng.uniform(-2, 2, (n, 3))
    p = rng.uniform(0.2, 5.0, n)
    E = p / 0.4 + 0.5 * rho * (u ** 2).sum(1)
    return np.column_stack([rho, rho[:, None] * u, E])


def random_normals(n, seed):
    v = np.random.default_rng(seed).standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def test_flux_matches_reference_to_an_ulp(oracle, ref):
    """Same formulas as EulerPhysicalFlux.cpp; the two compilers may contract a*b+c
    into an FMA differently, so entries agree to one rounding, not bit for bit."""
    for gk, gamma in ((0, 1.4), (1, 5.0 / 3.0)):
        for U in random_states(50, 1):
            a, b = oracle.flux(U, gamma), ref.flux(U, gk)
            assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()
        S = random_states(40, 2)
        for UL, UR in zip(S[:20], S[20:]):
            a, b = oracle.wave_speed(UL, UR, gamma), ref.wave_speed(UL, UR, gk)
            assert abs(a - b) <= 4.5e-16 * abs(b)


def test_flux_known_answer(oracle):
    # U = (1, 0, 0, 0, 2.5), gamma 1.4: p = 1, F = [0;I;0], |u|+c = sqrt(1.4)
    U = np.array([1.0, 0.0, 0.0, 0.0, 2.5])
    F = oracle.flux(U)
    expect = np.zeros((5, 3))
    expect[1:4] = np.eye(3)
    assert np.abs(F - expect).max() < 4e-16  # (gamma - 1) * 2.5 is 1 to a rounding
    assert abs(oracle.wave_speed(U, U) - np.sqrt(1.4)) < 1e-15
    # moving state: rho 2, u (1, -2, 0.5), p 3
    rho, u, p, g = 2.0, np.array([1.0, -2.0, 0.5]), 3.0, 1.4
    E = p / (g - 1) + 0.5 * rho * u @ u
    F = oracle.flux(np.concatenate([[rho], rho * u, [E]]))
    assert np.allclose(F[0], rho * u, rtol=1e-15)
    assert np.allclose(F[1:4], rho * np.outer(u, u) + p * np.eye(3), rtol=1e-14)
    assert np.allclose(F[4], u * (E + p), rtol=1e-14)


def test_density_clamp(oracle):
    # rho <= 0 is replaced by 1e-16 inside the flux (EulerPhysicalFlux.cpp:58-61)
    U = np.array([-1.0, 1e-17, 0.0, 0.0, 1e-16])
    F = oracle.flux(U)
    assert np.isfinite(F).all()


def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    return q * np.sign(np.linalg.det(q))


@pytest.mark.parametrize("kind", ["lf", "roe"])
def test_numflux_consistency_conservation_rotation(oracle, kind):
    S = random_states(60, 3)
    N = random_normals(30, 4)
    for i in range(30):
        UL, UR, n = S[i], S[30 + i], N[i]
        # consistency: Fhat(U, U, n) = F(U) n
        assert rel_err(oracle.numflux(kind, UL, UL, n), oracle.flux(UL) @ n) < 1e-13
        # conservation: Fhat(UL, UR, n) = -Fhat(UR, UL, -n)
        a = oracle.numflux(kind, UL, UR, n)
        b = oracle.numflux(kind, UR, UL, -n)
        assert rel_err(a, -b) < 1e-13
        # rotational invariance: Fhat(T UL, T UR, Q n) = T Fhat(UL, UR, n), T = diag(1, Q, 1)
        Q = rotation(i)
        T = np.eye(5)
        T[1:4, 1:4] = Q
        c = oracle.numflux(kind, T @ UL, T @ UR, Q @ n)
        assert rel_err(c, T @ a) < 1e-12


def test_lf_closed_form(oracle):
    S = random_states(20, 5)
    N = random_normals(10, 6)
    for i in range(10):
        UL, UR, n = S[i], S[10 + i], N[i]
        lam = oracle.wave_speed(UL, UR)
        expect = 0.5 * (oracle.flux(UL) @ n + oracle.flux(UR) @ n + lam * (UL - UR))
        assert rel_err(oracle.numflux("lf", UL, UR, n), expect) < 1e-14


def test_roe_property(oracle):
    """sum_k lambda_k alpha_k r_k over the Roe-averaged eigensystem equals (F(UR) - F(UL)) n:
    pins the wave strengths, eigenvectors and Roe averages (Roe 1981)."""
    S = random_states(40, 7)
    N = random_normals(20, 8)
    for i in range(20):
        UL, UR, n = S[i], S[20 + i], N[i]
        d = oracle.numflux("roe", UL, UR, n, signed_eigs=True)
        assert rel_err(d, (oracle.flux(UR) - oracle.flux(UL)) @ n) < 1e-12


def test_roe_upwinds_supersonic_flow(oracle):
    # all characteristic speeds positive (beyond the entropy-fix band) -> Fhat = F(UL) n
    g = 1.4
    rho, p = 1.0, 1.0
    c = np.sqrt(g * p / rho)
    n = np.array([1.0, 0.0, 0.0])
    UL = np.array([rho, rho * 3 * c, 0.0, 0.0, p / (g - 1) + 0.5 * rho * (3 * c) ** 2])
    UR = np.array([1.1 * rho, 1.1 * rho * 3.1 * c, 0.0, 0.0, 1.2 * p / (g - 1) + 0.5 * 1.1 * rho * (3.1 * c) ** 2])
    assert rel_err(oracle.numflux("roe", UL, UR, n), oracle.flux(UL) @ n) < 1e-13
    # and downwind for the mirrored case
    UL2, UR2 = UR.copy(), UL.copy()
    UL2[1], UR2[1] = -UL2[1], -UR2[1]
    assert rel_err(oracle.numflux("roe", UL2, UR2, n), oracle.flux(UR2) @ n) < 1e-13


def test_roe_resolves_stationary_contact_exactly(oracle):
    # density jump at rest with equal pressure: exact flux (0, p n, 0)
    g, p = 1.4, 2.0
    n = np.array([0.0, 0.6, 0.8])
    UL = np.array([1.0, 0, 0, 0, p / (g - 1)])
    UR = np.array([3.0, 0, 0, 0, p / (g - 1)])
    out = oracle.numflux("roe", UL, UR, n)
    assert np.allclose(out, np.concatenate([[0.0], p * n, [0.0]]), atol=1e-14)


def test_inverse3_matches_reference(oracle, ref):
    rng = np.random.default_rng(9)
    for _ in range(20):
        A = rng.standard_normal(9)
        a, b = oracle.inverse3(A), ref.inverse3(A)
        assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()
        assert np.abs(oracle.inverse3(A) @ A.reshape(3, 3) - np.eye(3)).max() < 1e-12


@pytest.mark.parametrize("bad", ["negative_energy", "negative_density"])
def test_roe_non_physical_state_takes_lax_friedrichs(oracle, bad):
    """A trace with p <= 0 or rho <= 0 has no real Roe-averaged sound speed; such points
    use the Lax-Friedrichs flux (the state below is a P3 face trace met on a coarse mesh)."""
    UL = np.array([0.5412775815615278, 0.22413491212667186, 1.0954546706020096, 0.0, 1.4438236016248314])
    UR = np.array([0.27407240678777567, -0.4779972781645934, 1.167592766300288, 0.0, -0.07210789244984943])
    if bad == "negative_density":
        UR = np.array([-0.01, 0.0, 0.1, 0.0, 1.0])
    n = np.array([0.0, -0.4472135954999579, 0.8944271909999159])
    roe = oracle.numflux("roe", UL, UR, n)
    assert np.isfinite(roe).all()
    assert np.array_equal(roe, oracle.numflux("lf", UL, UR, n))
    # physical states keep the Roe flux
    UR2 = np.array([0.6, 0.1, 1.0, 0.0, 2.0])
    assert not np.allclose(oracle.numflux("roe", UL, UR2, n), oracle.numflux("lf", UL, UR2, n))


def test_lf_dissipation_sign_and_size(oracle):
    """The Lax-Friedrichs dissipation against the jump: (U_L - U_R) . (Fhat - (F_L + F_R) n / 2)
    is positive and lies between 1/2 max(|u.n| + c) |U_L - U_R|^2 and 1/2 max(|u| + c) |U_L - U_R|^2
    (the sound speeds from the states by numpy): a flipped sign, a missing 1/2 or a normal-speed
    lambda fails one of the bounds."""
    g = 1.4
    S = random_states(40, 9)
    N = random_normals(20, 10)

    def speeds(U, n):
        rho, m, E = U[0], U[1:4], U[4]
        u = m / rho
        c = np.sqrt(g * (g - 1) * (E - 0.5 * rho * u @ u) / rho)
        return abs(u @ n) + c, np.linalg.norm(u) + c

    for i in range(20):
        UL, UR, n = S[i], S[20 + i], N[i]
        d = UL - UR
        diss = d @ (oracle.numflux("lf", UL, UR, n) - 0.5 * (oracle.flux(UL) @ n + oracle.flux(UR) @ n))
        (aL, bL), (aR, bR) = speeds(UL, n), speeds(UR, n)
        lo, hi = 0.5 * max(aL, aR) * (d @ d), 0.5 * max(bL, bR) * (d @ d)
        assert diss > 0
        assert lo * (1 - 1e-12) <= diss <= hi * (1 + 1e-12)
