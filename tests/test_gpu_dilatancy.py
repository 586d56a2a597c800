"""Dilatancy (assemble(gf, p, dila, u0, tspan), equation.jl:108-117,
168-183, 248-282) on the device against the oracle: the RHS for the
reference's own random-property case (test/BEM/tests.jl:73-81) and for a
physical one, with the dense and the Toeplitz operator; a short trajectory
against the oracle's Tsit5 over the oracle's RHS."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import relerr, rs_props, small_fault

pytestmark = pytest.mark.gpu
YR = 365 * 86400.0


def _physical(mf, rng):
    shp = (mf.nx, mf.nxi)
    return oq.DilatancyProperty(10 ** rng.uniform(5, 7, shp), 10 ** rng.uniform(-5, -3, shp),
                                10 ** rng.uniform(-10, -9, shp), rng.uniform(1e6, 5e6, shp))


@pytest.mark.parametrize("case,dip,kdtype", [("random", 90.0, "fp64"), ("physical", 45.0, "fp64"),
                                             ("physical", 90.0, "toeplitz"), ("physical", 60.0, "fft")])
def test_dilatancy_rhs_matches_oracle(oracle, rng, case, dip, kdtype):
    mf = small_fault(dip)
    pf = rs_props(mf, rng)
    shp = (mf.nx, mf.nxi)
    if case == "random":  # DilatancyProperty([rand(nx, nξ) for _ in 1:4]...)
        dila = oq.DilatancyProperty(*[rng.random(shp) for _ in range(4)])
        u = oq.ArrayPartition(*[rng.random(shp) for _ in range(4)])
    else:
        dila = _physical(mf, rng)
        v = pf.vpl * np.exp(rng.standard_normal(shp))
        u = oq.ArrayPartition(v, pf.L / v * np.exp(0.3 * rng.standard_normal(shp)), rng.random(shp),
                              dila.p0 * (1 + 0.3 * rng.standard_normal(shp)))
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    prob = oq.assemble(st, pf, dila, u, (0.0, 1.0), kdtype=kdtype)
    du = oq.similar(u)
    prob.f(du, u, prob.p, 1.0)
    ref = oracle.rhs_fault_dila(pf, dila, st, u.x[0], u.x[1], u.x[3])
    assert all(np.all(np.isfinite(x)) for x in du.x)
    for k, rk in zip((0, 1, 2, 3), ref):
        assert relerr(du.x[k], rk) <= 1e-12, k
    prob.f.close()


def test_dilatancy_trajectory_matches_oracle(oracle, rng):
    mf = small_fault(90.0)
    pf = rs_props(mf, rng)
    dila = _physical(mf, rng)
    shp = (mf.nx, mf.nxi)
    v = pf.vpl * np.exp(0.2 * rng.standard_normal(shp))
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros(shp), dila.p0.copy())
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    tspan = (0.0, 0.2 * YR)
    prob = oq.assemble(st, pf, dila, u0, tspan)
    sol = oq.solve(prob, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    assert sol.retcode == "Success"
    nf = mf.nx * mf.nxi

    def f(t, uf):
        parts = [uf[k * nf:(k + 1) * nf].reshape(shp, order="F") for k in range(4)]
        out = oracle.rhs_fault_dila(pf, dila, st, parts[0], parts[1], parts[3])
        return np.concatenate([o.reshape(-1, order="F") for o in out])

    uref, stats = oracle.tsit5(f, u0.pack(), tspan[0], tspan[1], reltol=1e-6, abstol=1e-8)
    assert stats.retcode == 0
    assert abs(sol.naccept - stats.naccept) <= 1
    got = sol.u.pack()
    for k in range(4):
        assert relerr(got[k * nf:(k + 1) * nf], uref[k * nf:(k + 1) * nf]) <= 1e-6, k


def test_dilatancy_argument_checks(oq, oracle, rng):
    mf = small_fault(90.0)
    pf = rs_props(mf, rng)
    dila = _physical(mf, rng)
    shp = (mf.nx, mf.nxi)
    u = oq.ArrayPartition(*[np.ones(shp) for _ in range(4)])
    st = oracle.gf_ff(mf, 3e10, 3e10)
    prob = oq.assemble(st, pf, dila, u, (0.0, 1.0))
    it = oq.Integrator(prob)
    with pytest.raises(ValueError):
        it.set_persistent(True)  # the persistent stepper has no dilatancy law
    it.close()
