"""wsolve: integrate on the device while streaming snapshots to disk
(src/io.jl:3-133).

The reference appends every `stride`-th accepted state (through `getu`)
to chunked, extensible HDF5 datasets with time as the trailing dimension,
flushing every `nstep` snapshots.  No HDF5 library is available here, so a
"file" is a directory store with the same datasets: `<name>.f64` (raw
little-endian float64, one snapshot after another) and `index.json`
(shapes and counts).  `wread(file, name)` returns a dataset like
`h5read(file, name)`: shape (size(u)..., nt), Julia (column-major) order.
The snapshots are taken on the device by the stepper's recorder.
"""
import json
import logging
import os

import numpy as np

from .equation import Integrator, ODESolution

log = logging.getLogger("oetqf_amd")


class _Store:
    def __init__(self, path, create):
        self.path = path
        self.ix = os.path.join(path, "index.json")
        if create:
            os.makedirs(path, exist_ok=True)
            for f in os.listdir(path):
                if f.endswith(".f64") or f == "index.json":
                    os.remove(os.path.join(path, f))
            self.meta = {}
        else:
            with open(self.ix) as f:
                self.meta = json.load(f)

    def append(self, name, arr):
        a = np.asarray(arr, dtype="<f8")
        shape = list(a.shape)
        m = self.meta.setdefault(name, {"shape": shape, "count": 0})
        if m["shape"] != shape:
            raise ValueError("shape of %s changed from %s to %s" % (name, m["shape"], shape))
        with open(os.path.join(self.path, name + ".f64"), "ab") as f:
            f.write(np.asfortranarray(a).tobytes(order="F"))
        m["count"] += 1

    def count(self, name):
        return self.meta.get(name, {"count": 0})["count"]

    def flush(self):
        tmp = self.ix + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.meta, f)
        os.replace(tmp, self.ix)


def wread(file, name):
    """The dataset `name` of a wsolve store: shape (size(u)..., nt) (h5read)."""
    with open(os.path.join(file, "index.json")) as f:
        m = json.load(f)[name]
    raw = np.fromfile(os.path.join(file, name + ".f64"), dtype="<f8")
    shape = tuple(m["shape"]) + (m["count"],)
    return raw.reshape(shape, order="F")


class _ValType:
    def __init__(self, k):
        self.k = k

    def __repr__(self):
        return "Val[%d]" % self.k


class Val:
    """Val[1] stands for Julia's Val{1}: integrator(t, Val[1]) asks for the
    first derivative of the solution at t."""

    _cache = {}

    def __class_getitem__(cls, k):
        return cls._cache.setdefault(int(k), _ValType(int(k)))


class SavedState:
    """The `integrator` argument that getu(u, t, integrator) receives in
    wsolve: the state saved at time `t`, and f(u) there when the recorder
    kept it.  integrator(integrator.t, Val[1]) returns the derivative as an
    ArrayPartition (the reference's handler 𝐕𝚯𝚬′𝚫 reads its third part, the
    viscous strain rate, src/io.jl:86)."""

    def __init__(self, t, u, du):
        self.t, self.u, self._du = t, u, du
        self.derivative_used = False

    def __call__(self, t, deriv=None):
        k = 0 if deriv is None else deriv.k
        if t != self.t:
            raise NotImplementedError("the device recorder keeps the saved state only: "
                                      "integrator(t, ...) needs t == integrator.t")
        if k == 0:
            return self.u
        if k == 1:
            self.derivative_used = True
            if self._du is None:
                raise RuntimeError("the derivative was not recorded for this state")
            return self._du
        raise NotImplementedError("only Val[0] and Val[1] are recorded")


def VΘΔ(u, *args):
    """Saving handler (src/io.jl:85): slip rate, state variable, slip of a
    fault-only problem."""
    return (u.x[0], u.x[1], u.x[2])


def VΘdεΔ(u, t, integrator):
    """Saving handler (src/io.jl:86, 𝐕𝚯𝚬′𝚫): slip rate, state variable, the
    viscous strain rate integrator(t, Val{1}).x[3] and slip of a coupled
    problem."""
    return (u.x[0], u.x[1], integrator(integrator.t, Val[1]).x[2], u.x[4])


def wsolve(prob, alg, file, nstep, getu, ustrs, tstr, stride=1, append=False, force=False,
           reltol=1e-6, abstol=1e-8, dt=0.0, dtmax=0.0, dtmin=0.0, maxiters=10**9,
           persistent=False):
    """wsolve(prob, alg, file, nstep, getu, ustrs, tstr; stride, append,
    force, kwargs...) (io.jl:118-133) with alg = Tsit5() or VCABM5().
    getu(u, t, integrator) is called on each saved state; `integrator` is a
    SavedState, so integrator(integrator.t, Val[1]) gives f(u) there (the
    derivative snapshots are recorded on the device when getu asks for them
    at t0).  Returns the final ODESolution, or None when the file exists and
    neither force nor append is set (io.jl:119-123).  persistent=True runs
    the stepper in one cooperative launch per drain (fault-only problems
    whose operator rows fit in LDS; Integrator.set_persistent)."""
    from .equation import Tsit5, VCABM5

    if not isinstance(alg, (Tsit5, VCABM5)):
        raise NotImplementedError("the device steppers are Tsit5() and VCABM5(); got %r" % (alg,))
    if os.path.exists(file) and not force and not append:
        log.info("Overwrite existing file %s must set `force = true`.", file)
        log.info("Aborting computation.")
        return None
    assert tstr not in ustrs, "Duplicate name of %s in %s." % (tstr, ustrs)
    t0 = prob.tspan[0]
    # getu's first call (io.jl:126) also tells whether it reads the derivative
    du0 = prob.u0.similar()
    prob.f(du0, prob.u0, prob.p, t0)
    first = SavedState(t0, prob.u0, du0)
    ptrs0 = getu(prob.u0, t0, first)
    assert len(ptrs0) == len(ustrs), "Unmatched length between solution components and names."
    want_du = first.derivative_used
    it = Integrator(prob, reltol=reltol, abstol=abstol, dt=dt, dtmax=dtmax, dtmin=dtmin,
                    maxiters=maxiters, alg=alg)
    try:
        if persistent:
            it.set_persistent(True)
        store = _Store(file, create=not append)

        def save(ptrs, t):
            assert len(ptrs) == len(ustrs), "Unmatched length between solution components and names."
            for name, a in zip(ustrs, ptrs):
                store.append(name, a)
            store.append(tstr, np.asarray(t, dtype=np.float64))

        # the callback's first call is at t0 (stride_count 0, io.jl:51-58)
        save(ptrs0, t0)
        it.record(stride=stride, snap_capacity=nstep, step_capacity=0, derivative=want_du)
        chunk = max(1, nstep)  # attempted steps per drain <= nstep snapshots
        ret = "Running"
        while ret == "Running":
            ret = it.step(chunk)
            out = it.drain()
            snap_t, snaps = out[2], out[3]
            dsnaps = out[4] if want_du else [None] * len(snap_t)
            for t, flat, dflat in zip(snap_t, snaps, dsnaps):
                u = prob.u0.similar().unpack(flat)
                du = prob.u0.similar().unpack(dflat) if dflat is not None else None
                save(getu(u, t, SavedState(t, u, du)), t)
            store.flush()
        store.flush()
        t, _, u, na, nr = it.state()
    finally:
        it.close()
    return ODESolution(t, u, ret, na, nr)
