"""Earthquake-cycle parity: the 512-patch strike-slip fault (40 x 20 km,
seeded V0) integrated with the on-device Tsit5 and with the oracle's Tsit5
over the same FP64 operator; the first two earthquakes (onsets ~16.4 and
~64.5 yr) must coincide to 1e-4 relative, as must their peak slip rates to
1e-2 (the peak is sampled at accepted steps)."""
import numpy as np
import pytest

import oetqf_amd as oq
from oetqf_amd.cycle import events, run_cycle
from oetqf_amd.workloads import YEAR, strike_slip_512

pytestmark = pytest.mark.gpu


def test_first_events_match_oracle(oracle):
    years = 70.0
    w = strike_slip_512()
    prob = oq.assemble_from_mesh(w["mf"], None, w["lam"], w["mu"], w["pf"], None, w["u0"],
                                 (0.0, years * YEAR))
    r = run_cycle(prob, reltol=1e-6, abstol=1e-8)
    assert r["retcode"] == "Success"
    eg = events(r["t"], r["vmax"], 1e-3, merge_gap=365 * 86400.0)
    K = prob.f.dense_operator()
    nf = w["mf"].nx * w["mf"].nxi
    _, st, tt, vv = oracle.cycle_dense(w["pf"], None, K, nf, 0, w["u0"].pack(), 0.0, years * YEAR,
                                       reltol=1e-6, abstol=1e-8)
    assert st.retcode == 0
    ec = events(tt, vv, 1e-3, merge_gap=365 * 86400.0)
    print("GPU onsets (yr):", [e["onset"] / YEAR for e in eg], "oracle:", [e["onset"] / YEAR for e in ec])
    assert len(eg) == len(ec) == 2
    for a, b in zip(eg, ec):
        assert abs(a["onset"] - b["onset"]) <= 1e-4 * b["onset"]
        assert abs(a["peak_v"] - b["peak_v"]) <= 1e-2 * b["peak_v"]
    assert abs(r["naccept"] - st.naccept) <= 0.01 * st.naccept
