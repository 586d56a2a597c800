"""Pins of the oracle's Tsit5 (oracle/tsit5.c).

OrdinaryDiffEq is not vendored (Project.toml:17), so the tableau is pinned
by the Runge-Kutta order conditions (Butcher trees up to order 5 for b,
order 4 for the embedded bhat = b - btilde) and the integrator by its
measured local order and by exact solutions.
"""
import numpy as np
import pytest


def test_tableau_order_conditions(oracle):
    c, A, b, bt = oracle.tsit5_tableau()
    assert np.allclose(A.sum(axis=1), c, atol=1e-14)
    bh = b - bt

    def conds(w, order):
        C = np.diag(c)
        out = [w.sum() - 1]
        if order >= 2:
            out += [w @ c - 1 / 2]
        if order >= 3:
            out += [w @ c**2 - 1 / 3, w @ A @ c - 1 / 6]
        if order >= 4:
            out += [w @ c**3 - 1 / 4, w @ C @ A @ c - 1 / 8, w @ A @ c**2 - 1 / 12, w @ A @ A @ c - 1 / 24]
        if order >= 5:
            Ac = A @ c
            out += [w @ c**4 - 1 / 5, w @ C @ C @ Ac - 1 / 10, w @ (Ac * Ac) - 1 / 20,
                    w @ C @ A @ c**2 - 1 / 15, w @ A @ c**3 - 1 / 20, w @ C @ A @ A @ c - 1 / 30,
                    w @ A @ C @ A @ c - 1 / 40, w @ A @ A @ c**2 - 1 / 60, w @ A @ A @ A @ c - 1 / 120]
        return np.abs(out)

    assert conds(b, 5).max() < 1e-12
    assert conds(bh, 4).max() < 1e-12
    assert conds(bh, 5).max() > 1e-6  # the embedded pair is genuinely 4th order
    assert np.allclose(A[6], b) and c[6] == 1.0  # FSAL


def test_local_order_is_five(oracle):
    """One step of y' = cos(t) y (y = exp(sin t)): local error O(h^6), so
    halving h divides it by ~2^6 (2^5 or 2^7 would mean a wrong order)."""
    f = lambda t, u: np.array([np.cos(t) * u[0]])
    errs = []
    for h in (0.1, 0.05):
        u, st = oracle.tsit5(f, [1.0], 0.0, h, reltol=1e3, abstol=1e3, dt0=h)
        assert st.naccept == 1
        errs.append(abs(u[0] - np.exp(np.sin(h))))
    assert 48 < errs[0] / errs[1] < 100


def test_adaptive_linear_system(oracle):
    lam = np.array([-1.0, -0.1, -0.01])
    f = lambda t, u: lam * u
    u, st = oracle.tsit5(f, [1.0, 2.0, 3.0], 0.0, 10.0, reltol=1e-8, abstol=1e-10)
    assert st.retcode == 0 and st.t == 10.0
    assert np.allclose(u, np.array([1.0, 2.0, 3.0]) * np.exp(lam * 10.0), rtol=1e-6)


def test_retcodes(oracle):
    f = lambda t, u: -u
    _, st = oracle.tsit5(f, [1.0], 0.0, 1e3, maxiters=3)
    assert st.retcode == 1 and st.naccept + st.nreject == 3
    _, st = oracle.tsit5(f, [1.0], 0.0, 1e3, dtmin=1e9)
    assert st.retcode == 2
    _, st = oracle.tsit5(lambda t, u: np.array([np.nan]), [1.0], 0.0, 1.0)
    assert st.retcode == 3
