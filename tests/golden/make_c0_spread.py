"""Writes tests/golden/c0_catalogue_spread.json: the oracle's 300-year event
catalogues of the 512-patch strike-slip fault under tiny perturbations of
the initial state.

The earthquake cycle of this fault is chaotic after its first two events
(~16.4 and ~64 yr): two exact FP64 integrations that differ only in the
summation order of the shear-stress sums follow different event sequences
from ~80 yr on.  A single 300-year catalogue therefore cannot be matched
event by event; its statistics can.  This script records, for Tsit5 and
VCABM5 and for seeds 0..23 (seed 0 = the unperturbed initial state; seed s >
0 multiplies theta_0 by 1 + 1e-9 N(0, 1) drawn from default_rng(s)), the
event count, the onsets, the peak slip rates and the step counts of the
oracle's integration over the oracle's own operator.  The device runs the
same seeded states and is compared with this spread.

Events: max slip rate above 1e-3 m/s, crossings within 1 year merged.
Run from the repository root (about 30 minutes on 8 cores):

    python tests/golden/make_c0_spread.py
"""
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = range(24)
ALGS = ("tsit5", "vcabm5")
PERTURB = 1e-9
MERGE_GAP = 365 * 86400.0


def _setup():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "oetqf.jl_amd"))


def perturbed_u0(w, seed):
    """The seeded initial state (shared with the device test)."""
    import numpy as np
    u0 = w["u0"].pack().copy()
    nf = w["mf"].nx * w["mf"].nxi
    if seed > 0:
        r = np.random.default_rng(seed)
        u0[nf:2 * nf] *= 1.0 + PERTURB * r.standard_normal(nf)
    return u0


def dense_kff(oracle, w):
    """The receiver-major dense K_ff of the oracle's Toeplitz kernel."""
    import numpy as np
    mf = w["mf"]
    st = oracle.gf_ff(mf, w["lam"], w["mu"], nrept=w["nrept"], buffer_ratio=w["buffer_ratio"])
    nf = mf.nx * mf.nxi
    K = np.zeros((nf, nf))
    for c in range(nf):
        e = np.zeros(nf)
        e[c] = 1.0
        K[:, c] = oracle.dtau_toeplitz(st, e.reshape(mf.nx, mf.nxi, order="F")).reshape(-1, order="F")
    return K


def _job(args):
    alg, seed = args
    os.environ["OMP_NUM_THREADS"] = "1"
    _setup()
    import numpy as np
    import oracle
    from oetqf_amd.cycle import events
    from oetqf_amd.workloads import YEAR, strike_slip_512
    w = strike_slip_512()
    K = dense_kff(oracle, w)
    nf = w["mf"].nx * w["mf"].nxi
    t0 = time.perf_counter()
    _, st, tt, vv = oracle.cycle_dense(w["pf"], None, K, nf, 0, perturbed_u0(w, seed), 0.0,
                                       300 * YEAR, reltol=w["reltol"], abstol=w["abstol"], alg=alg)
    ev = events(tt, vv, 1e-3, merge_gap=MERGE_GAP)
    on = np.array([e["onset"] for e in ev])
    return dict(alg=alg, seed=seed, retcode=int(st.retcode), naccept=int(st.naccept),
                nreject=int(st.nreject), events=len(ev), onsets_yr=(on / YEAR).tolist(),
                peak_v=[float(e["peak_v"]) for e in ev],
                mean_recurrence_yr=float(np.mean(np.diff(on)) / YEAR) if len(on) > 1 else None,
                wall_s=time.perf_counter() - t0)


def main():
    jobs = [(a, s) for a in ALGS for s in SEEDS]
    with mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        runs = pool.map(_job, jobs, chunksize=1)
    out = dict(workload="strike-slip fault, 512 patches (32x16), elastic, 300 yr",
               perturbation="theta_0 *= 1 + %g N(0,1), default_rng(seed), seed 0 unperturbed" % PERTURB,
               event_def="max v > 1e-3 m/s, crossings within 1 yr merged", runs=runs)
    with open(os.path.join(HERE, "c0_catalogue_spread.json"), "w") as f:
        json.dump(out, f, indent=1)
    for r in runs:
        print(r["alg"], r["seed"], r["events"], round(r["mean_recurrence_yr"], 3), r["naccept"],
              round(r["wall_s"], 1))


if __name__ == "__main__":
    main()
