"""Ensembles of frictional-parameter realizations sharing one geometry.

A parameter study over the reference means one `assemble` + `solve` per
realization (equation.jl:141-154).  `Ensemble` binds M members to one
FP32/BF16 operator on the device: their RHS is one MFMA GEMM and a lockstep
Tsit5 advances all of them with per-member step control.
"""
import ctypes

import numpy as np

from . import _lib
from .equation import _KDTYPES, _rs_struct, _visc_struct
from .greens import StrikeSlip, _Keep, _ftype, fault_struct, get_quadrature, hex8_struct, quad_struct


_ENS_KDTYPES = dict(_KDTYPES, fp16x2=_lib.OQ_K_FP16X2)


class Ensemble:
    def __init__(self, mf, ma, lam, mu, members, pa, ftype=StrikeSlip(), nrept=2, buffer_ratio=0.0,
                 qtype="Gauss1", kdtype="fp32", device=0):
        if kdtype not in ("fp32", "bf16", "fp16x2"):
            raise ValueError("the ensemble operator is fp32, bf16 or fp16x2")
        keep = _Keep()
        self._keep = keep
        M = len(members)
        arr = (_lib.RSProps * M)(*[_rs_struct(p, keep) for p in members])
        keep.append(arr)
        L = _lib.load()
        self.h = ctypes.c_void_p()
        _lib.register(self)
        a = fault_struct(mf, keep)
        if ma is not None:
            b = hex8_struct(ma, keep)
            q = quad_struct(get_quadrature(qtype), keep)
            vs = _visc_struct(pa, ma.ne, keep)
            _lib.check(L.oq_ensemble_create(device, ctypes.byref(a), ctypes.byref(b), lam, mu,
                                            _ftype(ftype), nrept, buffer_ratio, ctypes.byref(q), M,
                                            arr, ctypes.byref(vs), _ENS_KDTYPES[kdtype],
                                            ctypes.byref(self.h)))
            self.ne = ma.ne
        else:
            _lib.check(L.oq_ensemble_create(device, ctypes.byref(a), None, lam, mu, _ftype(ftype),
                                            nrept, buffer_ratio, None, M, arr, None,
                                            _ENS_KDTYPES[kdtype], ctypes.byref(self.h)))
            self.ne = 0
        self.M = M
        self._rec_cap = 0
        self.nf = mf.nx * mf.nxi
        self.n = 3 * self.nf + 12 * self.ne

    def eval(self, U):
        """U: (M, nstate) packed member states -> dU (M, nstate)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.M, self.n)
        dU = np.zeros_like(U)
        _lib.check(_lib.load().oq_ensemble_eval(self.h, _lib.ptr(U), _lib.ptr(dU)))
        return dU

    def start(self, U0, tspan, reltol=1e-6, abstol=1e-8, dt0=1.0, dtmax=0.0, dtmin=0.0,
              maxiters=10**9, alg=None):
        """Lockstep run of every member: Tsit5 (default) or alg=VCABM5()."""
        from .equation import Tsit5, VCABM5
        if alg is not None and not isinstance(alg, (Tsit5, VCABM5)):
            raise NotImplementedError("the ensemble steps with Tsit5() or VCABM5()")
        self._u0 = np.ascontiguousarray(U0, dtype=np.float64)
        o = _lib.SolverOpts(reltol, abstol, dt0, dtmax, dtmin, int(maxiters))
        fn = _lib.load().oq_ensemble_start_vcabm5 if isinstance(alg, VCABM5) else _lib.load().oq_ensemble_start
        _lib.check(fn(self.h, _lib.ptr(self._u0), float(tspan[0]), float(tspan[1]), ctypes.byref(o)))

    def record(self, step_capacity):
        """Per-member (t, max|v|) record of accepted steps; takes effect at
        the next start().  0 turns it off."""
        _lib.check(_lib.load().oq_ensemble_record(self.h, int(step_capacity)))
        self._rec_cap = int(step_capacity)

    def drain(self):
        """[(t, vmax)] per member: the accepted steps since the last drain."""
        cap = self._rec_cap
        cnt = np.zeros(self.M, dtype=np.int64)
        t = np.zeros((self.M, cap))
        v = np.zeros((self.M, cap))
        _lib.check_drain(_lib.load().oq_ensemble_drain(
            self.h, cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _lib.ptr(t), _lib.ptr(v)))
        return [(t[m, :cnt[m]].copy(), v[m, :cnt[m]].copy()) for m in range(self.M)]

    def checkpoint(self):
        """(U, ctrl) of the current run: the members' states and controller
        arrays (t, dt, qold, naccept, nreject, iter, done, retcode), for
        resume() here or in another process."""
        c = _lib.new_ctrl(self.M)
        _lib.check(_lib.load().oq_ensemble_get_ctrl(self.h, ctypes.byref(_lib.LockstepCtrl.of(c))))
        return self.state()[1], c

    def resume(self, U, ctrl, t1, reltol=1e-6, abstol=1e-8, dtmax=0.0, dtmin=0.0, maxiters=10**9):
        """Continues a checkpointed run towards t1 (bit for bit the
        uninterrupted run on the same operator)."""
        self._u0 = np.ascontiguousarray(U, dtype=np.float64)
        self._ctrl = {k: np.ascontiguousarray(v, dtype=_lib.new_ctrl(1)[k].dtype) for k, v in ctrl.items()}
        o = _lib.SolverOpts(reltol, abstol, 1.0, dtmax, dtmin, int(maxiters))
        _lib.check(_lib.load().oq_ensemble_resume(self.h, _lib.ptr(self._u0), ctypes.byref(_lib.LockstepCtrl.of(self._ctrl)),
                                  float(t1), ctypes.byref(o)))

    def run(self, max_steps):
        n = ctypes.c_int64()
        _lib.check(_lib.load().oq_ensemble_run(self.h, int(max_steps), ctypes.byref(n)))
        return n.value

    def state(self):
        t = np.zeros(self.M)
        U = np.zeros((self.M, self.n))
        na = np.zeros(self.M, dtype=np.int64)
        nr = np.zeros(self.M, dtype=np.int64)
        rc = np.zeros(self.M, dtype=np.int32)
        _lib.check(_lib.load().oq_ensemble_state(
            self.h, _lib.ptr(t), _lib.ptr(U), na.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            nr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            rc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return t, U, na, nr, [_lib.RETCODES[int(r)] for r in rc]

    def time(self, iters=10):
        ms = ctypes.c_double()
        _lib.check(_lib.load().oq_ensemble_time(self.h, int(iters), ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            _lib.load().oq_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
