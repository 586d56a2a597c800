"""Earthquake-cycle runs on the device and their event catalogue.

`run_cycle` drives the on-device Tsit5 through the recorder (one (t, max v)
record per accepted step); `events` turns that record into an earthquake
catalogue: an event starts when the largest slip rate on the fault crosses
`vthresh` upward and ends when it falls back below it.
"""
import time

import numpy as np

from .equation import Integrator


def run_cycle(prob, reltol=1e-6, abstol=1e-8, dt=0.0, dtmax=0.0, chunk=4096, wall_limit=None,
              progress=None, persistent=False, alg=None):
    """Integrates prob over its tspan on the device.  Returns a dict with the
    per-step record (t, vmax), the final state and counters.  persistent:
    True = one cooperative launch per chunk of steps (small fault-only
    problems; ValueError otherwise), "auto" = that where the problem
    qualifies, else the graph stepper."""
    it = Integrator(prob, reltol=reltol, abstol=abstol, dt=dt, dtmax=dtmax, alg=alg)
    it.record(stride=1, snap_capacity=0, step_capacity=chunk)
    mode = "graph"
    if persistent:
        try:
            it.set_persistent(True)
            mode = "persistent"
        except ValueError:
            if persistent != "auto":
                raise
    ts, vs = [], []
    t0 = time.perf_counter()
    ret = "Running"
    while ret == "Running":
        ret = it.step(chunk)
        st, sv, _, _ = it.drain()
        ts.append(st)
        vs.append(sv)
        if progress is not None:
            progress(st[-1] if len(st) else None)
        if wall_limit is not None and time.perf_counter() - t0 > wall_limit:
            break
    wall = time.perf_counter() - t0
    t, dtl, u, na, nr = it.state()
    it.close()
    return dict(t=np.concatenate(ts), vmax=np.concatenate(vs), retcode=ret, t_end=t, u=u,
                naccept=na, nreject=nr, wall=wall, mode=mode)


def events(step_t, step_vmax, vthresh=1e-3, merge_gap=0.0):
    """Catalogue of events from a (t, max v) record: onset time (linear
    interpolation of log v at the upward crossing), peak slip rate and the
    time of the peak.  Crossings that start less than `merge_gap` seconds
    after the previous event ended belong to that event (a rupture that
    flickers around the threshold is one event)."""
    v = np.asarray(step_vmax)
    t = np.asarray(step_t)
    above = v >= vthresh
    out = []
    n = len(v)
    if n < 2:
        return out
    # upward crossings (above[i], not above[i - 1]) and the first index below
    # the threshold after each of them
    starts = np.flatnonzero(above[1:] & ~above[:-1]) + 1
    falls = np.flatnonzero(~above[1:] & above[:-1]) + 1
    for i in starts:
        la, lb = np.log(v[i - 1]), np.log(v[i])
        f = (np.log(vthresh) - la) / (lb - la) if lb != la else 0.0
        onset = t[i - 1] + f * (t[i] - t[i - 1])
        q = np.searchsorted(falls, i)
        j = int(falls[q]) if q < len(falls) else n
        k = i + int(np.argmax(v[i:j]))
        end = t[j - 1]
        if out and onset - out[-1]["end"] < merge_gap:
            ev = out[-1]
            if v[k] > ev["peak_v"]:
                ev["peak_v"], ev["t_peak"] = float(v[k]), float(t[k])
            ev["end"] = float(end)
        else:
            out.append(dict(onset=float(onset), peak_v=float(v[k]), t_peak=float(t[k]), end=float(end)))
    return out


def recurrence(evs):
    on = np.array([e["onset"] for e in evs])
    return np.diff(on)


def run_ensemble_cycle(ens, U0, tspan, reltol=1e-6, abstol=1e-8, dt0=1.0, dtmax=0.0, chunk=2048,
                       wall_limit=None, resume=None, progress=None, alg=None):
    """Integrates all members of an Ensemble over tspan with the per-member
    recorder on.  Returns per-member (t, vmax) records, final states and
    counters; `complete` is False when wall_limit stopped it, and then
    `checkpoint` = (U, ctrl) continues it: pass the returned dict as
    `resume` (its records are kept in front of the new ones)."""
    ens.record(chunk)
    ts = [[] for _ in range(ens.M)]
    vs = [[] for _ in range(ens.M)]
    iters0, wall0 = 0, 0.0
    if resume is not None:
        U, ctrl = resume["checkpoint"]
        ens.resume(U, ctrl, tspan[1], reltol=reltol, abstol=abstol, dtmax=dtmax)
        for m in range(ens.M):
            ts[m].append(resume["t"][m])
            vs[m].append(resume["vmax"][m])
        iters0, wall0 = resume["iters"], resume["wall"]
    else:
        ens.start(U0, tspan, reltol=reltol, abstol=abstol, dt0=dt0, dtmax=dtmax, alg=alg)
    t0 = time.perf_counter()
    iters = iters0
    complete = False
    while True:
        left = ens.run(chunk)
        iters += chunk
        for m, (t, v) in enumerate(ens.drain()):
            ts[m].append(t)
            vs[m].append(v)
        if progress is not None:
            progress(iters, left)
        if left == 0:
            complete = True
            break
        if wall_limit is not None and time.perf_counter() - t0 > wall_limit:
            break
    wall = wall0 + time.perf_counter() - t0
    t, U, na, nr, rc = ens.state()
    return dict(t=[np.concatenate(x) for x in ts], vmax=[np.concatenate(x) for x in vs],
                t_end=t, U=U, naccept=na, nreject=nr, retcode=rc, wall=wall, iters=iters,
                complete=complete, checkpoint=None if complete else ens.checkpoint())
