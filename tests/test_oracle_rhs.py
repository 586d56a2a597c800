"""Pins of the oracle's RHS (oracle/rhs.c) -- src/BEM/equation.jl.

- the Toeplitz-DFT K_ff product equals the dense Toeplitz contraction
  (the reference's own test, test/BEM/tests.jl:39-61, tightened);
- the viscosity law closed form (test/BEM/tests.jl:142-154);
- update_fault! is the quasi-dynamic balance d/dt[tau_f(v, theta) + eta v]
  = dtau/dt for the regularized rate-and-state friction
  tau_f = sigma a asinh(v/(2 v0) exp((f0 + b ln(v0 theta/L))/a)) with the
  aging law (Rubin & Ampuero 2005), checked against finite differences of
  tau_f rather than by re-evaluating the same formula;
- the coupled RHS equals the same RHS on the dense receiver-major operator
  (block layout and sign contracts: source = v - vpl, equation.jl:38).
"""
import numpy as np
import pytest

from helpers import relerr, rs_props, small_fault, small_mantle, visc_props


@pytest.mark.parametrize("ftype", [0, 1])
def test_fft_conv_equals_dense_toeplitz(oracle, rng, ftype):
    # test/BEM/tests.jl:43-59: 100 x 100 fault, 10 x 10 cells, dip 41, lambda = mu = 3e10
    mf = oracle_mesh(oracle, 100.0, 100.0, 10.0, 10.0, 41.0)
    st = oracle.gf_ff(mf, 3e10, 3e10, ftype=ftype)
    v = rng.random((mf.nx, mf.nxi))
    relv = v - 0.1
    got = oracle.dtau_fft(st, relv)
    nx, nxi = mf.nx, mf.nxi
    i, j, k, l = np.meshgrid(np.arange(nx), np.arange(nxi), np.arange(nx), np.arange(nxi), indexing="ij")
    gf3 = st[np.abs(i - k), j, l]
    want = np.einsum("ijkl,kl->ij", gf3, relv)
    assert relerr(got, want) <= 1e-12
    assert relerr(oracle.dtau_toeplitz(st, relv), want) <= 1e-13


class _M:
    def __init__(self, d):
        self.__dict__.update(d)


def oracle_mesh(oracle, *a):
    return _M(oracle.mesh_rect_okada(*a))


class _Visc:
    def __init__(self, gamma, n, deps0):
        self.gamma, self.n, self.deps0 = np.asarray(gamma, float), np.asarray(n, float), np.asarray(deps0, float)


def test_viscosity_law_closed_form(oracle, rng):
    """dϵ_dt == γ σ τ^n (test/BEM/tests.jl:142-149) on the strain rate the
    oracle derives from a stress tensor: for a pure-shear stress sxy the
    second invariant tau = sqrt(2) sxy."""
    A, n = rng.random(), 2.0 * rng.random()
    ne = 3
    sig = np.zeros((ne, 6))
    sxy = rng.random(ne)
    sig[:, 1] = sxy
    de = oracle.update_strain_rate(_Visc(A * np.ones(ne), n * np.ones(ne), np.zeros(6)), sig)
    tau = np.sqrt(2.0) * sxy
    assert np.allclose(de[:, 1], A * sxy * tau**n, rtol=1e-14, atol=0)
    assert np.all(de[:, [0, 2, 3, 4, 5]] == 0)
    # isotropic stress has no deviator: no flow
    iso = np.zeros((ne, 6))
    iso[:, [0, 3, 5]] = 5.0
    assert np.all(oracle.update_strain_rate(_Visc(np.ones(ne), np.zeros(ne), np.zeros(6)), iso) == 0)


def _tau_f(p, i, v, th):
    a, b, L, s = p.a.flat[i], p.b.flat[i], p.L.flat[i], p.σ.flat[i]
    return s * a * np.arcsinh(v / (2 * p.v0) * np.exp((p.f0 + b * np.log(p.v0 * th / L)) / a))


def test_update_fault_is_quasi_dynamic_balance(oracle, oq, rng):
    mf = small_fault(90.0)
    p = rs_props(mf, rng)
    nf = mf.nx * mf.nxi
    v = p.vpl * np.exp(2 * rng.standard_normal(nf))
    th = p.L.reshape(-1, order="F") / v * np.exp(rng.standard_normal(nf))
    dtau = 1e-3 * rng.standard_normal(nf)
    dv, dth, dd = oracle.update_fault(p, v, th, dtau)
    a_f = np.asarray(p.a).reshape(-1, order="F")
    for i in range(0, nf, 7):
        pi = oq.RateStateQuasiDynamicProperty(*(np.asarray(x).reshape(-1, order="F")[i:i + 1]
                                                for x in (p.a, p.b, p.L, p.σ)), p.η, p.vpl)
        # step: a small fraction of the time scales of v and theta, large enough
        # that the rounding of tau_f (~1e7 Pa) does not swamp the rate
        h = 1e-3 * min(abs(v[i] / dv[i]), abs(th[i] / dth[i]))
        f = lambda s: _tau_f(pi, 0, v[i] + s * dv[i], th[i] + s * dth[i]) + p.η * (v[i] + s * dv[i])
        rate = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
        assert rate == pytest.approx(dtau[i], rel=1e-6, abs=1e-9 * abs(dtau).max())
        assert dth[i] == pytest.approx(1 - v[i] * th[i] / np.asarray(p.L).reshape(-1, order="F")[i], rel=1e-15)
        assert a_f[i] > 0
    assert np.array_equal(dd, v)


def test_steady_state_is_a_fixed_point(oracle, rng):
    """v = vpl, theta = L/v, no stressing: dtheta = 0 and dv = 0."""
    mf = small_fault(90.0)
    p = rs_props(mf, rng)
    nf = mf.nx * mf.nxi
    v = np.full(nf, p.vpl)
    th = np.asarray(p.L).reshape(-1, order="F") / v
    dv, dth, _ = oracle.update_fault(p, v, th, np.zeros(nf))
    assert np.abs(dth).max() <= 4e-16 and np.abs(dv).max() <= 1e-30


def test_nonpositive_theta_gives_radiation_damping_only(oracle, rng):
    """max(0, θ) (equation.jl:238): θ <= 0 gives ψ1 = 0, so dv = dτ/η."""
    mf = small_fault(90.0)
    p = rs_props(mf, rng)
    nf = mf.nx * mf.nxi
    v = np.full(nf, 1e-6)
    th = -np.ones(nf)
    dtau = rng.standard_normal(nf)
    dv, _, _ = oracle.update_fault(p, v, th, dtau)
    assert np.allclose(dv, dtau / p.η, rtol=1e-15, atol=0)


def test_coupled_rhs_equals_dense_operator_rhs(oracle, rng):
    mf, ma = small_fault(90.0), small_mantle(2, 2, 1)
    lam = mu = 3e10
    st = oracle.gf_ff(mf, lam, mu, buffer_ratio=1.0)
    g12 = oracle.gf_mf(mf, ma, lam, mu, buffer_ratio=1.0)
    g21 = oracle.gf_fm(ma, mf, lam, mu)
    g22 = oracle.gf_mm(ma, lam, mu)
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 1.5)
    nx, nxi, ne = mf.nx, mf.nxi, ma.ne
    nf = nx * nxi
    v = pf.vpl * np.exp(rng.standard_normal((nx, nxi)))
    th = pf.L / v
    sig = 1e6 * rng.standard_normal((ne, 6))
    ref = oracle.rhs_coupled(pf, pa, st, g12, g21, g22, v, th, sig)
    i, j, k, l = np.meshgrid(np.arange(nx), np.arange(nxi), np.arange(nx), np.arange(nxi), indexing="ij")
    kff = st[np.abs(i - k), j, l].transpose(1, 0, 3, 2).reshape(nf, nf)
    K = np.block([[kff, g21], [g12, g22]])
    got = oracle.rhs_dense(pf, pa, np.ascontiguousarray(K), nf, ne, v.reshape(-1, order="F"),
                           th.reshape(-1, order="F"), sig.reshape(-1, order="F"))
    assert relerr(got[0], ref[0].reshape(-1, order="F")) <= 1e-12
    assert relerr(got[2], ref[2].reshape(-1, order="F")) <= 1e-15
    assert relerr(got[3], ref[3].reshape(-1, order="F")) <= 1e-12
    # sign contract: the source is v - vpl (equation.jl:38): v = vpl everywhere
    # and no viscous flow give no stressing
    z = oracle.rhs_coupled(pf, visc_props(ma, rng, 0.0).__class__(np.zeros(ne), np.zeros(ne), np.zeros(6)),
                           st, g12, g21, g22, np.full((nx, nxi), pf.vpl), th, sig)
    assert np.abs(z[3]).max() == 0.0


def _dila_props(mf, rng):
    from types import SimpleNamespace
    shp = (mf.nx, mf.nxi)
    return SimpleNamespace(tp=10 ** rng.uniform(5, 7, shp), eps=10 ** rng.uniform(-5, -3, shp),
                           beta=10 ** rng.uniform(-10, -9, shp), p0=rng.uniform(1e6, 5e6, shp))


def _tau_f_dila(a, b, L, sigma, f0, v0, v, th, pp):
    """Power-law rate-and-state friction on the effective normal stress
    (Segall & Rice, the form of equation.jl:248-276): f0 (sigma - p)
    (v/v0)^(a/f0) (theta v0/L)^(b/f0)."""
    return f0 * (sigma - pp) * max(0.0, v / v0) ** (a / f0) * max(0.0, th * v0 / L) ** (b / f0)


def test_dilatancy_update_is_quasi_static_balance(oracle, rng):
    """The rates of v, theta and p keep the friction equal to the stress:
    d/dt tau_f(v, theta, p) = dtau (no radiation damping in this law); theta
    follows the aging law; delta' = v."""
    mf = small_fault(90.0)
    p = rs_props(mf, rng)
    d = _dila_props(mf, rng)
    nf = mf.nx * mf.nxi
    flat = lambda x: np.asarray(x).reshape(-1, order="F")
    v = p.vpl * np.exp(2 * rng.standard_normal(nf))
    th = flat(p.L) / v * np.exp(rng.standard_normal(nf))
    pp = flat(d.p0) * (1 + 0.5 * rng.standard_normal(nf))
    dtau = 1e-3 * rng.standard_normal(nf)
    dv, dth, dd, dp = oracle.update_fault_dila(p, d, v, th, pp, dtau)
    a, b, L, sg = flat(p.a), flat(p.b), flat(p.L), flat(p.σ)
    for i in range(0, nf, 5):
        h = 1e-3 * min(abs(v[i] / dv[i]), abs(th[i] / dth[i]), abs(pp[i] / dp[i]))
        f = lambda s: _tau_f_dila(a[i], b[i], L[i], sg[i], p.f0, p.v0, v[i] + s * dv[i],
                                  th[i] + s * dth[i], pp[i] + s * dp[i])
        rate = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
        assert rate == pytest.approx(dtau[i], rel=1e-6, abs=1e-9 * abs(dtau).max())
        assert dth[i] == pytest.approx(1 - v[i] * th[i] / L[i], rel=1e-15)
    assert np.array_equal(dd, v)


def test_dilatancy_pressure_law(oracle, rng):
    """At a steady state variable the pore pressure relaxes linearly toward
    p0 on the time scale tp; at p = p0 it follows (eps/beta) d(ln theta)/dt."""
    mf = small_fault(90.0)
    p = rs_props(mf, rng)
    d = _dila_props(mf, rng)
    nf = mf.nx * mf.nxi
    flat = lambda x: np.asarray(x).reshape(-1, order="F")
    v = p.vpl * np.exp(rng.standard_normal(nf))
    th_ss = flat(p.L) / v
    p0, tp = flat(d.p0), flat(d.tp)
    z = np.zeros(nf)
    _, dth, _, dp1 = oracle.update_fault_dila(p, d, v, th_ss, p0 + 1e5, z)
    _, _, _, dp2 = oracle.update_fault_dila(p, d, v, th_ss, p0 + 3e5, z)
    assert np.abs(dth).max() <= 4e-16
    assert np.allclose(dp1, -1e5 / tp, rtol=1e-12) and np.allclose(dp2, 3 * dp1, rtol=1e-12)
    th = th_ss * np.exp(rng.standard_normal(nf))
    _, dth, _, dp = oracle.update_fault_dila(p, d, v, th, p0, z)
    ratio = flat(d.eps) / flat(d.beta)
    for i in range(0, nf, 5):
        h = 1e-4 * th[i] / max(abs(dth[i]), 1e-300)
        g = lambda s: ratio[i] * np.log(th[i] + s * dth[i])
        assert (g(h) - g(-h)) / (2 * h) == pytest.approx(dp[i], rel=1e-6)


def test_dilatancy_rhs_uses_fft_traction_rate(oracle, rng):
    """The dilatancy ode (equation.jl:168-183) = FFT traction rate + the
    pointwise update."""
    mf = small_fault(60.0)
    p = rs_props(mf, rng)
    d = _dila_props(mf, rng)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    v = p.vpl * np.exp(rng.standard_normal((mf.nx, mf.nxi)))
    th = np.asarray(p.L) / v
    pp = np.asarray(d.p0) * 1.1
    out = oracle.rhs_fault_dila(p, d, st, v, th, pp)
    dtau = oracle.dtau_fft(st, v - p.vpl)
    ref = oracle.update_fault_dila(p, d, v.reshape(-1, order="F"), th.reshape(-1, order="F"),
                                   pp.reshape(-1, order="F"), np.asarray(dtau).reshape(-1, order="F"))
    for k in range(4):
        assert np.array_equal(out[k].reshape(-1, order="F"), ref[k])
