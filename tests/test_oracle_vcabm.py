"""The oracle's VCABM5 (variable-coefficient Adams-Bashforth-Moulton, order
5, PECE; oracle/vcabm.c) pinned by the mathematics of the method (Hairer,
Norsett & Wanner I, Sec. III.5): the coefficients against the classical
constant-step Adams coefficients and against the defining integrals for
variable steps, the predictor's exactness on polynomials through divided
differences, the divided-difference update, and convergence on a smooth
problem.  Its step sequence against OrdinaryDiffEq's VCABM5 is unpinned
(OrdinaryDiffEq is not vendored in the reference)."""
from fractions import Fraction

import numpy as np
import pytest


def _psi(h, hist):
    hist = list(hist) + [0.0] * 6
    new = [h + sum(hist[:i]) for i in range(6)]      # t_{n+1} - t_{n-i}
    old = [sum(hist[:i + 1]) for i in range(6)]      # t_n - t_{n-i-1}
    return new, old


def test_constant_steps_give_adams_bashforth_coefficients(oracle):
    # gamma_j of the explicit Adams methods: 1, 1/2, 5/12, 3/8, 251/720, 95/288
    gam = [Fraction(1), Fraction(1, 2), Fraction(5, 12), Fraction(3, 8), Fraction(251, 720),
           Fraction(95, 288)]
    g, beta = oracle.vcabm_coeffs(5, 0.3, [0.3] * 5)
    assert np.allclose(g, [float(x) for x in gam], rtol=1e-14, atol=0)
    assert np.allclose(beta, 1.0, rtol=1e-15)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_variable_step_coefficients_equal_defining_integrals(oracle, seed):
    rng = np.random.default_rng(seed)
    hist = list(rng.uniform(0.2, 2.0, 5))
    h = rng.uniform(0.2, 2.0)
    g, beta = oracle.vcabm_coeffs(5, h, hist)
    new, old = _psi(h, hist)
    x, w = np.polynomial.legendre.leggauss(12)
    s = 0.5 * h * (x + 1.0)
    for j in range(6):
        integrand = np.ones_like(s)
        for i in range(j):  # (t - t_{n-i}) / (t_{n+1} - t_{n-i}),  t - t_n = s
            integrand *= (s + (new[i] - h)) / new[i]
        gj = 0.5 * np.sum(w * integrand)
        assert g[j] == pytest.approx(gj, rel=1e-13), j
    for j in range(5):
        assert beta[j] == pytest.approx(np.prod([new[i] / old[i] for i in range(j)]), rel=1e-14)


def _phi(ts, fs, j):
    """Modified divided difference Phi_j(n) = prod_{i<j}(t_n - t_{n-i-1})
    f[t_n, ..., t_{n-j}] with ts[0] = t_n, ts[1] = t_{n-1}, ..."""
    d = list(fs[:j + 1])
    for lev in range(1, j + 1):
        d = [(d[i] - d[i + 1]) / (ts[i] - ts[i + lev]) for i in range(len(d) - 1)]
    return np.prod([ts[0] - ts[i + 1] for i in range(j)]) * d[0]


def test_predictor_integrates_quartics_exactly(oracle):
    """h sum_j g_j beta_j Phi_j(n) = int_{t_n}^{t_{n+1}} P for a polynomial P
    of degree <= 4 sampled at five unevenly spaced points; and the update
    Phi_{j+1}(n+1) = Phi_j(n+1) - beta_j Phi_j(n) gives the divided
    differences at n+1."""
    rng = np.random.default_rng(7)
    c = rng.standard_normal(5)
    P = np.polynomial.Polynomial(c)
    hist = list(rng.uniform(0.3, 1.5, 5))
    h = 0.9
    tn = 10.0
    ts = [tn - sum(hist[:i]) for i in range(6)]  # t_n, t_{n-1}, ...
    fs = [P(t) for t in ts]
    g, beta = oracle.vcabm_coeffs(5, h, hist)
    phi = [_phi(ts, fs, j) for j in range(5)]
    pred = h * sum(g[j] * beta[j] * phi[j] for j in range(5))
    exact = P.integ()(tn + h) - P.integ()(tn)
    assert pred == pytest.approx(exact, rel=1e-12)
    ts1, fs1 = [tn + h] + ts, [P(tn + h)] + fs
    tmp = fs1[0]
    for j in range(5):
        assert tmp == pytest.approx(_phi(ts1, fs1, j), rel=1e-10, abs=1e-10)
        tmp = tmp - beta[j] * phi[j]


def test_vcabm5_converges_on_smooth_problem(oracle):
    """y' = cos(t) y, y(0) = 1: y = exp(sin t)."""
    f = lambda t, y: np.cos(t) * y
    errs = []
    for tol in (1e-6, 1e-9):
        u, st = oracle.vcabm5(f, np.array([1.0]), 0.0, 10.0, reltol=tol, abstol=tol)
        assert st.retcode == 0
        errs.append(abs(u[0] - np.exp(np.sin(10.0))))
        assert errs[-1] <= 50 * tol
    assert errs[1] < errs[0] / 100


def test_vcabm5_first_step_is_trapezoid(oracle):
    """k = 1: Euler predictor, trapezoid corrector; one step of y' = -y."""
    u, st = oracle.vcabm5(lambda t, y: -y, np.array([1.0]), 0.0, 0.1, reltol=1e-1, abstol=1e-1,
                          dt0=0.1)
    assert st.naccept == 1
    p = 1.0 - 0.1
    assert u[0] == pytest.approx(p + 0.05 * (-p - (-1.0)), rel=1e-15)
