"""The persistent stepper (one cooperative launch per chunk of steps, rows
of K resident in LDS, one grid barrier per RHS) against the graph stepper:
same controller and method (Tsit5, or VCABM5 with its divided differences
kept in LDS), so the same accepted steps and the same state up to FP64
summation order; the recorder gives the same catalogue."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import relerr, rs_props, small_fault, small_mantle, visc_props
from oetqf_amd.cycle import events, run_cycle
from oetqf_amd.workloads import YEAR, strike_slip_512

pytestmark = pytest.mark.gpu


def _run(prob, persistent, chunk, alg=None):
    it = oq.Integrator(prob, reltol=1e-6, abstol=1e-8, alg=alg)
    it.record(stride=7, snap_capacity=chunk, step_capacity=chunk)
    if persistent:
        it.set_persistent(True)
    ts, snaps = [], []
    ret = "Running"
    while ret == "Running":
        ret = it.step(chunk)
        st, sv, sn_t, sn_u = it.drain()
        ts.append(st)
        snaps.append(sn_u)
    out = it.state()
    it.close()
    return ret, out, np.concatenate(ts), np.concatenate(snaps)


@pytest.mark.parametrize("alg", ["tsit5", "vcabm5"])
@pytest.mark.parametrize("dip", [90.0, 45.0])
def test_persistent_matches_graph_small(oracle, rng, dip, alg):
    mf = small_fault(dip)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(0.3 * rng.standard_normal((mf.nx, mf.nxi)))
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros_like(v))
    prob = oq.assemble(st, pf, u0, (0.0, 0.5 * YEAR))
    a = oq.VCABM5() if alg == "vcabm5" else None
    rg, (tg, _, ug, nag, nrg), stg, sng = _run(prob, False, 300, a)
    rp, (tp, _, up, nap, nrp), stp, snp = _run(prob, True, 300, a)
    assert rg == rp == "Success" and tg == tp
    assert abs(nag - nap) <= 1 and abs(nrg - nrp) <= 1
    assert relerr(up.pack(), ug.pack()) <= 1e-8
    k = min(len(stg), len(stp))
    assert np.allclose(stp[:k], stg[:k], rtol=1e-6)
    assert sng.shape[1] == snp.shape[1] and abs(len(sng) - len(snp)) <= 1


@pytest.mark.parametrize("alg", ["tsit5", "vcabm5"])
def test_persistent_cycle_512(oq, alg):
    """Two earthquakes of the 512-patch fault (70 years): the persistent run
    gives the graph run's catalogue."""
    w = strike_slip_512()
    prob = oq.assemble_from_mesh(w["mf"], None, w["lam"], w["mu"], w["pf"], None, w["u0"],
                                 (0.0, 70 * YEAR))
    a = oq.VCABM5() if alg == "vcabm5" else None
    rg = run_cycle(prob, persistent=False, alg=a)
    rp = run_cycle(prob, persistent=True, alg=a)
    print("%s: graph %.2f s, persistent %.2f s, steps %d / %d" % (alg, rg["wall"], rp["wall"], rg["naccept"],
                                                                 rp["naccept"]), flush=True)
    assert rp["retcode"] == "Success"
    eg = events(rg["t"], rg["vmax"], 1e-3, merge_gap=YEAR)
    ep = events(rp["t"], rp["vmax"], 1e-3, merge_gap=YEAR)
    assert len(eg) == len(ep) == 2
    for a, b in zip(eg, ep):
        assert abs(a["onset"] - b["onset"]) <= 1e-6 * b["onset"]
        assert abs(a["peak_v"] - b["peak_v"]) <= 1e-3 * b["peak_v"]
    assert abs(rg["naccept"] - rp["naccept"]) <= 0.001 * rg["naccept"]


def test_persistent_rejects_coupled(oq, oracle, rng):
    mf, ma = small_fault(90.0), small_mantle()
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 0.0)
    v = pf.vpl * np.ones((mf.nx, mf.nxi))
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros((ma.ne, 6)), np.zeros((ma.ne, 6)), np.zeros_like(v))
    prob = oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u0, (0.0, 1e6))
    it = oq.Integrator(prob)
    with pytest.raises(ValueError):
        it.set_persistent(True)
    it.close()
