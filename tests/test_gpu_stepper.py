"""On-device Tsit5 vs the oracle's Tsit5 (same controller) on the same RHS.

Reference: solve(prob, Tsit5(); reltol, abstol) (test/tests.jl:11;
examples/otf-with-mantle.jl:160-161).
"""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import fault_state, relerr, rs_props, small_fault, small_mantle, visc_props

pytestmark = pytest.mark.gpu
YR = 365 * 86400.0


def _fault_problem(oracle, rng):
    mf = small_fault(90.0)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(0.1 * rng.standard_normal((mf.nx, mf.nxi)))
    th = pf.L / v
    return mf, st, pf, (v, th, np.zeros_like(v))


def test_stepper_matches_oracle_short_horizon(oq, oracle, rng):
    mf, st, pf, (v, th, dl) = _fault_problem(oracle, rng)
    u0 = oq.ArrayPartition(v, th, dl)
    tspan = (0.0, 0.5 * YR)
    prob = oq.assemble(st, pf, u0, tspan)
    sol = oq.solve(prob, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    assert sol.retcode == "Success"
    nf = mf.nx * mf.nxi

    def f(t, uf):
        vv = uf[:nf].reshape(mf.nx, mf.nxi, order="F")
        tt = uf[nf:2 * nf].reshape(mf.nx, mf.nxi, order="F")
        dv, dth, dd = oracle.rhs_fault(pf, st, vv, tt)
        return np.concatenate([dv.reshape(-1, order="F"), dth.reshape(-1, order="F"),
                               dd.reshape(-1, order="F")])

    uref, stats = oracle.tsit5(f, u0.pack(), tspan[0], tspan[1], reltol=1e-6, abstol=1e-8)
    assert stats.retcode == 0
    assert abs(sol.naccept - stats.naccept) <= 1
    got = sol.u.pack()
    for k in range(3):
        assert relerr(got[k * nf:(k + 1) * nf], uref[k * nf:(k + 1) * nf]) <= 1e-6


def test_stepper_coupled_runs(oq, oracle, rng):
    mf, ma = small_fault(90.0), small_mantle()
    pf, pa = rs_props(mf, rng), visc_props(ma, rng)
    v, th, dl = fault_state(mf, pf, rng)
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros((ma.ne, 6)), 1e5 * np.ones((ma.ne, 6)), dl)
    prob = oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u0, (0.0, 0.1 * YR), buffer_ratio=1.0)
    sol = oq.solve(prob, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    assert sol.retcode == "Success"
    assert sol.t == pytest.approx(0.1 * YR, rel=0, abs=0)
    assert all(np.all(np.isfinite(a)) for a in sol.u.x)


def test_stepper_retcodes(oq, oracle, rng):
    """MaxIters and DtLessThanMin forced from the host (the reference's
    solver options maxiters / dtmin)."""
    mf, st, pf, (v, th, dl) = _fault_problem(oracle, rng)
    u0 = oq.ArrayPartition(v, th, dl)
    prob = oq.assemble(st, pf, u0, (0.0, 300 * YR))
    sol = oq.solve(prob, oq.Tsit5(), maxiters=5)
    assert sol.retcode == "MaxIters" and sol.naccept + sol.nreject == 5
    sol = oq.solve(prob, oq.Tsit5(), dtmin=1e30)
    assert sol.retcode == "DtLessThanMin"
