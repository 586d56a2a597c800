"""Ensemble earthquake cycles against the FP64 oracle: three frictional
realizations (a - b, L, sigma drawn as for the 256-member study) on the
512-patch strike-slip fault advance together in the lockstep ensemble
stepper, with K streamed in FP32 through the FP32 MFMA GEMM or as the
split-FP16 operator on the FP16 matrix cores.  Each member's
catalogue over 80 years must match the oracle's FP64 Tsit5 run of the same
member: same event count, onsets within 1e-3 relative, recurrence intervals
within 1 %, peak slip rates within 5 %."""
import numpy as np
import pytest

import oetqf_amd as oq
from oetqf_amd.cycle import events, recurrence, run_ensemble_cycle
from oetqf_amd.workloads import YEAR, ensemble_members, strike_slip_512

pytestmark = pytest.mark.gpu
PICK = (0, 5, 6)  # members of ensemble_members(w, 8) with 1-3 events in 80 years
YEARS = 80.0
_ORACLE = {}  # member -> the oracle's (stats, step_t, step_vmax), shared by both operator types


@pytest.mark.parametrize("kdtype", ["fp32", "fp16x2"])
def test_ensemble_cycle_matches_oracle(oracle, kdtype):
    w = strike_slip_512()
    members, U0 = ensemble_members(w, 8)
    members = [members[m] for m in PICK]
    U0 = U0[list(PICK)]
    ens = oq.Ensemble(w["mf"], None, w["lam"], w["mu"], members, None, kdtype=kdtype)
    r = run_ensemble_cycle(ens, U0, (0.0, YEARS * YEAR), reltol=1e-6, abstol=1e-8, dt0=1.0)
    ens.close()
    assert r["retcode"] == ["Success"] * len(PICK)
    prob = oq.assemble_from_mesh(w["mf"], None, w["lam"], w["mu"], w["pf"], None, w["u0"],
                                 (0.0, YEARS * YEAR))
    K = prob.f.dense_operator()
    prob.f.close()
    nf = w["mf"].nx * w["mf"].nxi
    nev = 0
    for j, p in enumerate(members):
        assert len(r["t"][j]) == r["naccept"][j]
        eg = events(r["t"][j], r["vmax"][j], 1e-3, merge_gap=YEAR)
        if PICK[j] not in _ORACLE:
            _, st, tt, vv = oracle.cycle_dense(p, None, K, nf, 0, U0[j], 0.0, YEARS * YEAR,
                                               reltol=1e-6, abstol=1e-8, dt0=1.0)
            _ORACLE[PICK[j]] = (st, tt, vv)
        st, tt, vv = _ORACLE[PICK[j]]
        assert st.retcode == 0
        ec = events(tt, vv, 1e-3, merge_gap=YEAR)
        print("%s member %d: GPU onsets" % (kdtype, PICK[j]), ["%.4f" % (e["onset"] / YEAR) for e in eg],
              "oracle", ["%.4f" % (e["onset"] / YEAR) for e in ec],
              "naccept", r["naccept"][j], st.naccept, flush=True)
        assert len(eg) == len(ec)
        nev += len(ec)
        for a, b in zip(eg, ec):
            assert abs(a["onset"] - b["onset"]) <= 1e-3 * b["onset"]
            assert abs(a["peak_v"] - b["peak_v"]) <= 5e-2 * b["peak_v"]
        if len(ec) > 1:
            rg, rc = recurrence(eg), recurrence(ec)
            assert np.all(np.abs(rg - rc) <= 1e-2 * rc)
        assert abs(r["naccept"][j] - st.naccept) <= 0.02 * st.naccept
    assert nev >= 4
