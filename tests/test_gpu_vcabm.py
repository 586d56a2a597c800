"""VCABM5 on the device (oq_stepper_create_vcabm5) against the oracle's
VCABM5 (oracle/vcabm.c, the same method): short trajectories of the fault
and the coupled model, and the two earthquakes of the 512-patch fault."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import fault_state, relerr, rs_props, small_fault, small_mantle, visc_props
from oetqf_amd.cycle import events, run_cycle
from oetqf_amd.workloads import YEAR, strike_slip_512

pytestmark = pytest.mark.gpu


def test_vcabm5_matches_oracle_fault(oracle, rng):
    mf = small_fault(60.0)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    pf = rs_props(mf, rng)
    v, th, dl = fault_state(mf, pf, rng)
    u0 = oq.ArrayPartition(v, th, dl)
    tspan = (0.0, 0.5 * YEAR)
    prob = oq.assemble(st, pf, u0, tspan)
    sol = oq.solve(prob, oq.VCABM5(), reltol=1e-6, abstol=1e-8)
    assert sol.retcode == "Success"
    nf = mf.nx * mf.nxi

    def f(t, uf):
        vv = uf[:nf].reshape(mf.nx, mf.nxi, order="F")
        tt = uf[nf:2 * nf].reshape(mf.nx, mf.nxi, order="F")
        out = oracle.rhs_fault(pf, st, vv, tt)
        return np.concatenate([o.reshape(-1, order="F") for o in out])

    uref, stats = oracle.vcabm5(f, u0.pack(), tspan[0], tspan[1], reltol=1e-6, abstol=1e-8)
    assert stats.retcode == 0
    print("VCABM5 steps: device %d+%d, oracle %d+%d" % (sol.naccept, sol.nreject, stats.naccept,
                                                       stats.nreject))
    assert abs(sol.naccept - stats.naccept) <= 1
    got = sol.u.pack()
    for k in range(3):
        assert relerr(got[k * nf:(k + 1) * nf], uref[k * nf:(k + 1) * nf]) <= 1e-6


def test_vcabm5_matches_oracle_coupled(oracle, rng):
    mf, ma = small_fault(90.0), small_mantle()
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 0.0)
    v, th, dl = fault_state(mf, pf, rng)
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros((ma.ne, 6)), 1e5 * np.ones((ma.ne, 6)), dl)
    tspan = (0.0, 0.1 * YEAR)
    prob = oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u0, tspan, buffer_ratio=1.0)
    sol = oq.solve(prob, oq.VCABM5(), reltol=1e-6, abstol=1e-8)
    assert sol.retcode == "Success"
    K = prob.f.dense_operator()
    nf = mf.nx * mf.nxi
    uref, st, _, _ = oracle.cycle_dense(pf, pa, K, nf, ma.ne, u0.pack(), tspan[0], tspan[1],
                                        alg="vcabm5")
    assert st.retcode == 0 and abs(sol.naccept - st.naccept) <= 1
    assert relerr(sol.u.pack(), uref) <= 1e-6


def test_vcabm5_cycle_512(oracle):
    years = 70.0
    w = strike_slip_512()
    prob = oq.assemble_from_mesh(w["mf"], None, w["lam"], w["mu"], w["pf"], None, w["u0"],
                                 (0.0, years * YEAR))
    r = run_cycle(prob, alg=oq.VCABM5())
    rt = run_cycle(prob)
    assert r["retcode"] == "Success"
    eg = events(r["t"], r["vmax"], 1e-3, merge_gap=YEAR)
    K = prob.f.dense_operator()
    _, st, tt, vv = oracle.cycle_dense(w["pf"], None, K, 512, 0, w["u0"].pack(), 0.0, years * YEAR,
                                       alg="vcabm5")
    ec = events(tt, vv, 1e-3, merge_gap=YEAR)
    print("VCABM5 %.2f s (%d+%d steps, 2 RHS each), Tsit5 %.2f s (%d+%d steps, 6 RHS each); "
          "onsets %s vs oracle %s" % (r["wall"], r["naccept"], r["nreject"], rt["wall"], rt["naccept"],
                                      rt["nreject"], [round(e["onset"] / YEAR, 4) for e in eg],
                                      [round(e["onset"] / YEAR, 4) for e in ec]), flush=True)
    assert len(eg) == len(ec) == 2
    for a, b in zip(eg, ec):
        assert abs(a["onset"] - b["onset"]) <= 1e-4 * b["onset"]
        assert abs(a["peak_v"] - b["peak_v"]) <= 1e-2 * b["peak_v"]
    assert abs(r["naccept"] - st.naccept) <= 0.01 * st.naccept
