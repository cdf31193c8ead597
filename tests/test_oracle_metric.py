"""Pins the oracle's Kerr-Schild metric core and Hamiltonian RHS
(kerr_schild_core.hpp) with the reference's own checks
(test_geodesic_observables.cpp:7-66) and with facts the mathematics fixes
independently of the restated formulas: r solves
x^2+y^2+z^2 = r^2 + a^2 (1 - z^2/r^2) (x^2+y^2 = (r^2+a^2) sin^2 th, z = r cos th;
the header comment at kerr_schild_core.hpp:9 drops the a^2 term),
det g = -1, g g^-1 = I, the Boyer-Lindquist identity
g^{mu nu} d_mu r d_nu r = Delta / Sigma, the Schwarzschild limit, and the
analytic derivatives against 40-digit numerical differentiation (mpmath).
"""
import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")

A_REF = 1.0   # the reference tests run with the default globals a = M = 1
M_REF = 1.0

POINTS = [(0.0, 5.0, -7.0, 2.0), (0.0, 6.0, 2.0, 1.0), (0.0, -3.0, 4.5, -2.5),
          (0.0, 0.7, -0.4, 1.9), (0.0, 40.0, 13.0, -25.0)]


def test_reference_observables_cases(oracle):
    # test_geodesic_observables.cpp:7-27
    for x in (10.0, 12.0):
        obs = oracle.observables([0.0, x, 0.0, 0.0, -1.0, 1.0, 0.0, 0.0], A_REF, M_REF)
        assert abs(obs["energy"] - 1.0) < 1e-14
        assert abs(obs["angular_momentum_z"]) < 1e-14
        assert abs(obs["carter"]) < 1e-12
    # :29-36, hand-computed L_z = x p_y - y p_x
    obs = oracle.observables([0.0, 3.0, 4.0, 0.0, -2.0, 0.5, -0.3, 0.0], A_REF, M_REF)
    assert abs(obs["energy"] - 2.0) < 1e-14
    assert abs(obs["angular_momentum_z"] - (3.0 * -0.3 - 4.0 * 0.5)) < 1e-14


def test_reference_null_momentum_and_rhs(oracle):
    # test_geodesic_observables.cpp:38-66
    X = np.array([0.0, 5.0, -7.0, 2.0])
    ok, p = oracle.null_momentum(X, [0.4, -0.5, 0.768], A_REF, M_REF)
    assert ok
    obs = oracle.observables(np.concatenate([X, p]), A_REF, M_REF)
    assert abs(obs["norm"]) < 1e-10
    rhs = oracle.rhs(X, p, A_REF, M_REF)
    gi = oracle.inverse_metric(X, A_REF, M_REF)
    dgi = oracle.inverse_metric_deriv(X, A_REF, M_REF)
    for mu in range(4):
        assert abs(rhs[mu] - gi[mu] @ p) < 1e-12
        assert abs(rhs[4 + mu] + 0.5 * p @ dgi[mu] @ p) < 1e-12


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9, 1.0])
def test_radius_solves_spheroid_relation(oracle, a):
    for _, x, y, z in POINTS:
        r = oracle.lib().orc_compute_r(x, y, z, a)
        lhs = x * x + y * y + z * z
        rhs = r * r + a * a * (1.0 - z * z / (r * r))
        assert abs(lhs - rhs) <= 1e-13 * lhs


@pytest.mark.parametrize("a,m", [(0.0, 1.0), (0.5, 1.0), (0.9, 1.0), (1.0, 1.0), (0.3, 2.0)])
def test_metric_identities(oracle, a, m):
    for X in POINTS:
        X = np.array(X)
        g = oracle.metric(X, a, m)
        gi = oracle.inverse_metric(X, a, m)
        assert np.allclose(g, g.T, atol=0)
        assert abs(np.linalg.det(g) + 1.0) < 1e-11          # Kerr-Schild: det g = -1
        assert np.allclose(g @ gi, np.eye(4), atol=1e-12)
        # g^{mu nu} d_mu r d_nu r = Delta / Sigma (Boyer-Lindquist g^rr)
        rr = np.zeros(1)
        dr = np.zeros(3)
        oracle.lib().orc_compute_r_and_deriv(X[1], X[2], X[3], a, oracle.dptr(rr),
                                             oracle.dptr(dr))
        r = rr[0]
        grad = np.array([0.0, *dr])
        cos2 = (X[3] / r) ** 2
        delta = r * r - 2 * m * r + a * a
        sigma = r * r + a * a * cos2
        assert abs(grad @ gi @ grad - delta / sigma) < 1e-12 * max(1.0, abs(delta / sigma))


def test_schwarzschild_limit(oracle):
    m = 1.3
    for X in POINTS:
        X = np.array(X)
        g = oracle.metric(X, 0.0, m)
        r = np.linalg.norm(X[1:])
        n = X[1:] / r
        assert abs(g[0, 0] + (1 - 2 * m / r)) < 1e-14
        assert np.allclose(g[1:, 1:], np.eye(3) + 2 * m / r * np.outer(n, n), atol=1e-14)
        assert np.allclose(g[0, 1:], 2 * m / r * n, atol=1e-14)


def _mp_inverse_metric(X, a, m):
    """g^{mu nu} from the Kerr-Schild closed form in 40-digit arithmetic."""
    x, y, z = (mpmath.mpf(v) for v in X[1:])
    a = mpmath.mpf(a)
    R2 = x * x + y * y + z * z
    A = R2 - a * a
    r = mpmath.sqrt((A + mpmath.sqrt(A * A + 4 * a * a * z * z)) / 2)
    B = r * r + a * a
    lu = [-1, (r * x + a * y) / B, (r * y - a * x) / B, z / r]
    H = m * r ** 3 / (r ** 4 + a * a * z * z)
    eta = [-1, 1, 1, 1]
    return [[(eta[i] if i == j else 0) - 2 * H * lu[i] * lu[j] for j in range(4)] for i in range(4)]


@pytest.mark.parametrize("a,m", [(0.9, 1.0), (1.0, 1.0), (0.5, 2.0)])
def test_rhs_against_numerical_derivative(oracle, a, m):
    """dp_k/dlambda = -1/2 d_k g^{ab} p_a p_b with d_k from 40-digit central
    differences of the closed-form inverse metric."""
    mpmath.mp.dps = 40
    for X in POINTS[:4]:
        if oracle.lib().orc_compute_r(X[1], X[2], X[3], a) < 2.2 * m:
            continue     # inside the ergoregion no static future-directed root exists
        ok, p = oracle.null_momentum(np.array(X), [0.3, -0.6, 0.5], a, m)
        assert ok
        rhs = oracle.rhs(np.array(X), p, a, m)
        P = [mpmath.mpf(v) for v in p]

        def ham(*q):
            gi = _mp_inverse_metric((0.0,) + q, a, m)
            return sum(gi[i][j] * P[i] * P[j] for i in range(4) for j in range(4)) / 2

        for k in range(3):
            args = [mpmath.mpf(v) for v in X[1:]]
            d = mpmath.diff(lambda t: ham(*[args[i] if i != k else t for i in range(3)]),
                            args[k])
            assert abs(float(-d) - rhs[5 + k]) < 1e-11 * max(1.0, abs(float(d)))
        gi = np.array([[float(v) for v in row] for row in _mp_inverse_metric(X, a, m)])
        assert np.allclose(rhs[:4], gi @ p, rtol=1e-13, atol=1e-13)
        assert rhs[4] == 0.0


def test_null_momentum_is_null_and_future_directed(oracle):
    rng = np.random.default_rng(7)
    for _ in range(50):
        X = np.array([0.0, *rng.uniform(-30, 30, 3)])
        K = rng.normal(size=3)
        a = rng.uniform(0, 1)
        r = oracle.lib().orc_compute_r(X[1], X[2], X[3], a)
        if r < 2.5:
            continue
        ok, p = oracle.null_momentum(X, K, a, 1.0)
        assert ok
        gi = oracle.inverse_metric(X, a, 1.0)
        assert abs(p @ gi @ p) < 1e-12 * (p @ p)
        Kc = gi @ p
        assert Kc[0] > 0 and np.allclose(Kc[1:], K, rtol=1e-12, atol=1e-12)
