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


def wsolve(prob, alg, file, nstep, getu, ustrs, tstr, stride=1, append=False, force=False,
           reltol=1e-6, abstol=1e-8, dt=0.0, dtmax=0.0, dtmin=0.0, maxiters=10**9):
    """wsolve(prob, alg, file, nstep, getu, ustrs, tstr; stride, append,
    force, kwargs...) (io.jl:118-133).  getu(u, t, integrator) is called on
    each saved state (integrator = None: derivatives are not available from
    the device recorder).  Returns the final ODESolution, or None when the
    file exists and neither force nor append is set (io.jl:119-123)."""
    if os.path.exists(file) and not force and not append:
        log.info("Overwrite existing file %s must set `force = true`.", file)
        log.info("Aborting computation.")
        return None
    assert tstr not in ustrs, "Duplicate name of %s in %s." % (tstr, ustrs)
    it = Integrator(prob, reltol=reltol, abstol=abstol, dt=dt, dtmax=dtmax, dtmin=dtmin,
                    maxiters=maxiters)
    store = _Store(file, create=not append)
    t0 = prob.tspan[0]

    def save(u, t):
        ptrs = getu(u, t, None)
        assert len(ptrs) == len(ustrs), "Unmatched length between solution components and names."
        for name, a in zip(ustrs, ptrs):
            store.append(name, a)
        store.append(tstr, np.asarray(t, dtype=np.float64))

    # the callback's first call is at t0 (stride_count 0, io.jl:51-58)
    save(prob.u0, t0)
    it.record(stride=stride, snap_capacity=nstep, step_capacity=0)
    chunk = max(1, nstep)  # attempted steps per drain <= nstep snapshots
    ret = "Running"
    while ret == "Running":
        ret = it.step(chunk)
        _, _, snap_t, snaps = it.drain()
        for t, flat in zip(snap_t, snaps):
            save(prob.u0.similar().unpack(flat), t)
        store.flush()
    store.flush()
    t, _, u, na, nr = it.state()
    it.close()
    return ODESolution(t, u, ret, na, nr)
