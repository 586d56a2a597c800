"""Grouped batches: geometries of uneven size, each with up to 16
frictional realizations, on one device.

A parameter study over several geometries means one `assemble` + `solve`
per (geometry, realization) in the reference (equation.jl:141-154).  `Group`
binds them to one object: each geometry's operator is assembled on the
device and stored in FP32, one RHS of every member is one grouped MFMA
launch (per tile shape in use), tile shapes are picked per size bucket by
`autotune`, and a lockstep Tsit5 advances all members with per-member step
control.
"""
import ctypes

import numpy as np

from . import _lib
from .equation import _KDTYPES, _rs_struct, _visc_struct
from .greens import StrikeSlip, _Keep, _ftype, fault_struct, get_quadrature, hex8_struct, quad_struct

SHAPES = ["RB1-pipe", "RB1", "RB2-pipe", "RB2", "RB4-pipe", "RB4", "RB1-lds", "RB2-lds",
          "RB1-lds-d2", "RB2-lds-d2", "RB2-lds-d2-split", "dma-r64-k128", "dma-r32-k128-2wg", "dma-r64-k64-2wg", "dma-r64-k64-s2-4wg"]


class Group:
    """geoms: list of (mf, ma or None, pa or None); members: list over
    geometries of lists of M RateStateQuasiDynamicProperty (same M)."""

    def __init__(self, geoms, lam, mu, members, ftype=StrikeSlip(), nrept=2, buffer_ratio=0.0,
                 qtype="Gauss1", kdtype="fp32", device=0):
        if kdtype != "fp32":
            raise ValueError("the grouped operator is fp32")
        G = len(geoms)
        M = len(members[0])
        if len(members) != G or any(len(m) != M for m in members):
            raise ValueError("members must hold the same number of realizations per geometry")
        keep = _Keep()
        self._keep = keep
        fa = (_lib.FaultMesh * G)(*[fault_struct(mf, keep) for mf, _, _ in geoms])
        hp = (ctypes.POINTER(_lib.Hex8Mesh) * G)()
        vp = (ctypes.POINTER(_lib.ViscProps) * G)()
        for g, (mf, ma, pa) in enumerate(geoms):
            if ma is not None:
                h = hex8_struct(ma, keep)
                v = _visc_struct(pa, ma.ne, keep)
                keep.append(h)
                keep.append(v)
                hp[g] = ctypes.pointer(h)
                vp[g] = ctypes.pointer(v)
        q = quad_struct(get_quadrature(qtype), keep)
        arr = (_lib.RSProps * (G * M))(*[_rs_struct(p, keep) for mem in members for p in mem])
        keep.append(arr)
        self.h = ctypes.c_void_p()
        _lib.register(self)
        _lib.check(_lib.load().oq_group_create(device, G, fa, hp, vp, lam, mu, _ftype(ftype), nrept,
                                               buffer_ratio, ctypes.byref(q), M, arr,
                                               _KDTYPES[kdtype], ctypes.byref(self.h)))
        self.G, self.M = G, M
        ntot = ctypes.c_int64()
        kb = ctypes.c_int64()
        self.nstate = np.zeros(G, dtype=np.int64)
        _lib.check(_lib.load().oq_group_info(self.h, ctypes.byref(ntot),
                                             self.nstate.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                             ctypes.byref(kb)))
        self.ntot, self.kbytes = ntot.value, kb.value
        self.N = [mf.nx * mf.nxi + (6 * ma.ne if ma is not None else 0) for mf, ma, _ in geoms]

    def pack(self, states):
        """states[g]: (M, nstate[g]) -> flat host layout."""
        return np.concatenate([np.asarray(s, dtype=np.float64).reshape(-1) for s in states])

    def unpack(self, flat):
        out, o = [], 0
        for n in self.nstate:
            out.append(flat[o:o + n * self.M].reshape(self.M, n))
            o += n * self.M
        return out

    def autotune(self, iters=5):
        """Times every tile shape per size bucket, then the whole grouped GEMM
        with that choice against each single shape for all buckets, and keeps
        the faster; returns {bucket: (shape, ms)} and keeps the per-bucket
        table in self.tune_table {bucket: {shape: ms}}."""
        sh = np.zeros(8, dtype=np.int32)
        ms = np.zeros(8)
        tab = np.zeros(8 * len(SHAPES))
        _lib.check(_lib.load().oq_group_autotune(self.h, int(iters),
                                                 sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                 _lib.ptr(ms), _lib.ptr(tab)))
        self.tune_table = {b: {SHAPES[s]: float(tab[b * len(SHAPES) + s]) for s in range(len(SHAPES))}
                           for b in range(8) if sh[b] >= 0}
        return {b: (SHAPES[sh[b]], float(ms[b])) for b in range(8) if sh[b] >= 0}

    def set_shapes(self, shapes):
        """Pins tile shapes: {bucket: shape name or index}; buckets left out
        keep their shape."""
        sh = np.full(8, -1, dtype=np.int32)
        for b, s in shapes.items():
            sh[int(b)] = SHAPES.index(s) if isinstance(s, str) else int(s)
        _lib.check(_lib.load().oq_group_set_shapes(self.h, sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def eval(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.ntot,)
        dU = np.zeros_like(U)
        _lib.check(_lib.load().oq_group_eval(self.h, _lib.ptr(U), _lib.ptr(dU)))
        return dU

    def time(self, iters=10):
        ms = ctypes.c_double()
        _lib.check(_lib.load().oq_group_time(self.h, int(iters), ctypes.byref(ms)))
        return ms.value

    def start(self, U0, tspan, reltol=1e-6, abstol=1e-8, dt0=1.0, dtmax=0.0, dtmin=0.0,
              maxiters=10**9, alg=None):
        """Lockstep run of every member: Tsit5 (default) or alg=VCABM5()."""
        from .equation import Tsit5, VCABM5
        if alg is not None and not isinstance(alg, (Tsit5, VCABM5)):
            raise NotImplementedError("the grouped batch steps with Tsit5() or VCABM5()")
        self._u0 = np.ascontiguousarray(U0, dtype=np.float64)
        o = _lib.SolverOpts(reltol, abstol, dt0, dtmax, dtmin, int(maxiters))
        fn = _lib.load().oq_group_start_vcabm5 if isinstance(alg, VCABM5) else _lib.load().oq_group_start
        _lib.check(fn(self.h, _lib.ptr(self._u0), float(tspan[0]), float(tspan[1]), ctypes.byref(o)))

    def checkpoint(self):
        """(U, ctrl) of the current run: the members' states and controller
        arrays (t, dt, qold, naccept, nreject, iter, done, retcode), for
        resume() here or in another process."""
        c = _lib.new_ctrl(self.G * self.M)
        _lib.check(_lib.load().oq_group_get_ctrl(self.h, ctypes.byref(_lib.LockstepCtrl.of(c))))
        return self.state()[1], c

    def resume(self, U, ctrl, t1, reltol=1e-6, abstol=1e-8, dtmax=0.0, dtmin=0.0, maxiters=10**9):
        """Continues a checkpointed run towards t1 (bit for bit the
        uninterrupted run on the same operator)."""
        self._u0 = np.ascontiguousarray(U, dtype=np.float64)
        self._ctrl = {k: np.ascontiguousarray(v, dtype=_lib.new_ctrl(1)[k].dtype) for k, v in ctrl.items()}
        o = _lib.SolverOpts(reltol, abstol, 1.0, dtmax, dtmin, int(maxiters))
        _lib.check(_lib.load().oq_group_resume(self.h, _lib.ptr(self._u0), ctypes.byref(_lib.LockstepCtrl.of(self._ctrl)),
                                  float(t1), ctypes.byref(o)))

    def run(self, max_steps):
        n = ctypes.c_int64()
        _lib.check(_lib.load().oq_group_run(self.h, int(max_steps), ctypes.byref(n)))
        return n.value

    def state(self):
        GM = self.G * self.M
        t = np.zeros(GM)
        U = np.zeros(self.ntot)
        na = np.zeros(GM, dtype=np.int64)
        nr = np.zeros(GM, dtype=np.int64)
        rc = np.zeros(GM, dtype=np.int32)
        _lib.check(_lib.load().oq_group_state(
            self.h, _lib.ptr(t), _lib.ptr(U), na.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            nr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            rc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return t, U, na, nr, [_lib.RETCODES[int(r)] for r in rc]

    def close(self):
        if self.h:
            _lib.load().oq_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
