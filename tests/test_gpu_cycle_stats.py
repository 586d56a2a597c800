"""300-year event statistics of the 512-patch strike-slip fault against the
oracle's spread (tests/golden/c0_catalogue_spread.json).

The cycle is chaotic after its first events: a 1e-9 relative perturbation
of theta_0 changes the oracle's own 300-year catalogue (21-24 events).  The
device runs the same seeded initial states with the persistent steppers and
must (i) reproduce, for each seed, the events that precede the divergence
(the first three Tsit5 onsets, the first VCABM5 onset, within 1e-4
relative), and (ii) fall inside the oracle's distribution: every event count
within the oracle's range over its 24 seeds widened by one, and the mean
recurrence over the device seeds within three standard errors of the
oracle's mean."""
import json
import os

import numpy as np
import pytest

import oetqf_amd as oq
from golden.make_c0_spread import MERGE_GAP, perturbed_u0
from oetqf_amd.cycle import events, run_cycle
from oetqf_amd.workloads import YEAR, strike_slip_512

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = {"tsit5": (0, 1), "vcabm5": (0, 1, 2, 3)}
SHADOW = {"tsit5": 3, "vcabm5": 1}  # events before the oracle's seeds diverge


@pytest.mark.parametrize("alg", ["tsit5", "vcabm5"])
def test_c0_event_statistics_within_oracle_spread(oq, alg):
    fx = json.load(open(os.path.join(HERE, "golden", "c0_catalogue_spread.json")))
    ref = {r["seed"]: r for r in fx["runs"] if r["alg"] == alg}
    counts = np.array([r["events"] for r in ref.values()])
    recs = np.array([r["mean_recurrence_yr"] for r in ref.values()])
    w = strike_slip_512()
    got_n, got_rec = [], []
    for seed in SEEDS[alg]:
        u0 = w["u0"].similar().unpack(perturbed_u0(w, seed))
        prob = oq.assemble_from_mesh(w["mf"], None, w["lam"], w["mu"], w["pf"], None, u0, w["tspan"])
        r = run_cycle(prob, reltol=w["reltol"], abstol=w["abstol"], persistent=True,
                      alg=oq.VCABM5() if alg == "vcabm5" else None)
        prob.f.close()
        assert r["retcode"] == "Success"
        ev = events(r["t"], r["vmax"], 1e-3, merge_gap=MERGE_GAP)
        on = np.array([e["onset"] for e in ev]) / YEAR
        print("%s seed %d: %d events, mean recurrence %.3f yr (oracle %d, %.3f); %.1f s" %
              (alg, seed, len(ev), np.diff(on).mean(), ref[seed]["events"],
               ref[seed]["mean_recurrence_yr"], r["wall"]), flush=True)
        k = SHADOW[alg]
        assert np.allclose(on[:k], ref[seed]["onsets_yr"][:k], rtol=1e-4, atol=0.0)
        got_n.append(len(ev))
        got_rec.append(np.diff(on).mean())
    assert all(counts.min() - 1 <= n <= counts.max() + 1 for n in got_n), (got_n, counts)
    se = np.sqrt(recs.var(ddof=1) / len(recs) + recs.var(ddof=1) / len(got_rec))
    assert abs(np.mean(got_rec) - recs.mean()) <= 3 * se, (np.mean(got_rec), recs.mean(), se)
