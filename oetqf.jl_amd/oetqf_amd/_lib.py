"""ctypes binding of lib/liboetqf_amd.so (C ABI in include/oetqf_amd.h).

The shared library is the product: every Green's function, RHS evaluation
and time step below runs in its HIP kernels.  There is no CPU fallback --
if the library is missing or no GPU is visible, calls raise.
"""
import atexit
import ctypes
import os
import warnings
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(_PKG, "lib", "liboetqf_amd.so")

_d = ctypes.c_double
_i64 = ctypes.c_int64
_int = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

OQ_OK = 0
OQ_ERR_ARG = 1
OQ_ERR_STATE = 3
OQ_K_FP64, OQ_K_FP32, OQ_K_BF16, OQ_K_TOEPLITZ, OQ_K_FP16X2, OQ_K_FFT, OQ_K_FFT_GRID = 0, 1, 2, 3, 4, 5, 6
RETCODES = {0: "Running", 1: "Success", 2: "MaxIters", 3: "DtLessThanMin", 4: "NonFinite"}


class FaultMesh(ctypes.Structure):
    _fields_ = [("x", _pd), ("ax", _pd), ("y", _pd), ("z", _pd), ("axi", _pd),
                ("nx", _i64), ("nxi", _i64), ("dx", _d), ("dep", _d), ("dip", _d)]


class Hex8Mesh(ctypes.Structure):
    _fields_ = [("cx", _pd), ("cy", _pd), ("cz", _pd), ("qx", _pd), ("qy", _pd),
                ("qz", _pd), ("dx", _pd), ("dy", _pd), ("dz", _pd), ("ne", _i64)]


class Quadrature(ctypes.Structure):
    _fields_ = [("coords", _pd), ("weights", _pd), ("nq", _i64)]


class RSProps(ctypes.Structure):
    _fields_ = [("a", _pd), ("b", _pd), ("L", _pd), ("sigma", _pd),
                ("eta", _d), ("vpl", _d), ("f0", _d), ("v0", _d)]


class ViscProps(ctypes.Structure):
    _fields_ = [("gamma", _pd), ("n", _pd), ("nlaws", _i64), ("deps0", _d * 6)]


class DilaProps(ctypes.Structure):
    _fields_ = [("tp", _pd), ("eps", _pd), ("beta", _pd), ("p0", _pd)]


class SolverOpts(ctypes.Structure):
    _fields_ = [("reltol", _d), ("abstol", _d), ("dt0", _d), ("dtmax", _d), ("dtmin", _d),
                ("maxiters", _i64)]


class OetqfError(RuntimeError):
    """A non-zero status from the C ABI."""


_lib = None

# name -> (restype, argtypes); the exported surface of include/oetqf_amd.h
class LockstepCtrl(ctypes.Structure):
    """oq_lockstep_ctrl: per-member controller arrays of a lockstep run."""
    _fields_ = [("t", ctypes.POINTER(ctypes.c_double)), ("dt", ctypes.POINTER(ctypes.c_double)),
                ("qold", ctypes.POINTER(ctypes.c_double)), ("naccept", ctypes.POINTER(ctypes.c_int64)),
                ("nreject", ctypes.POINTER(ctypes.c_int64)), ("iter", ctypes.POINTER(ctypes.c_int64)),
                ("done", ctypes.POINTER(ctypes.c_int)), ("retcode", ctypes.POINTER(ctypes.c_int))]

    @classmethod
    def of(cls, c):
        """The structure over the numpy arrays of a controller dict (kept
        alive by the caller)."""
        pd, pl, pi = (ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                      ctypes.POINTER(ctypes.c_int))
        return cls(*[c[k].ctypes.data_as(p) for k, p in (("t", pd), ("dt", pd), ("qold", pd), ("naccept", pl),
                                                          ("nreject", pl), ("iter", pl), ("done", pi),
                                                          ("retcode", pi))])


def new_ctrl(n):
    """Empty controller arrays of n members."""
    return dict(t=np.zeros(n), dt=np.zeros(n), qold=np.zeros(n), naccept=np.zeros(n, np.int64),
                nreject=np.zeros(n, np.int64), iter=np.zeros(n, np.int64), done=np.zeros(n, np.int32),
                retcode=np.zeros(n, np.int32))


PROTOTYPES = {
    "oq_last_error": (ctypes.c_char_p, []),
    "oq_version": (_int, []),
    "oq_device_count": (_int, [ctypes.POINTER(_int)]),
    "oq_gf_okada_ff": (_int, [ctypes.POINTER(FaultMesh), _d, _d, _int, _i64, _d, _pd]),
    "oq_gf_okada_ff_dft": (_int, [ctypes.POINTER(FaultMesh), _d, _d, _int, _i64, _d, _pd]),
    "oq_gf_okada_ff_dft_noalloc": (_int, [ctypes.POINTER(FaultMesh), _d, _d, _int, _i64, _d, _vp, _vp]),
    "oq_gf_okada_mf": (_int, [ctypes.POINTER(FaultMesh), ctypes.POINTER(Hex8Mesh), _d, _d, _int,
                              ctypes.POINTER(Quadrature), _i64, _d, _pd]),
    "oq_gf_hex8_fm": (_int, [ctypes.POINTER(Hex8Mesh), ctypes.POINTER(FaultMesh), _d, _d, _int,
                             _pd]),
    "oq_gf_hex8_mm": (_int, [ctypes.POINTER(Hex8Mesh), _d, _d, ctypes.POINTER(Quadrature), _pd]),
    "oq_rhs_create": (_int, [_int, _i64, _i64, _i64, ctypes.POINTER(RSProps),
                             ctypes.POINTER(ViscProps), _pd, _int, _pd, _pd, _pd, _int,
                             ctypes.POINTER(_vp)]),
    "oq_rhs_create_from_mesh": (_int, [_int, ctypes.POINTER(FaultMesh), ctypes.POINTER(Hex8Mesh),
                                       _d, _d, _int, _i64, _d, ctypes.POINTER(Quadrature),
                                       ctypes.POINTER(RSProps), ctypes.POINTER(ViscProps), _int,
                                       ctypes.POINTER(_vp)]),
    "oq_rhs_eval": (_int, [_vp, _d] + [_pd] * 10),
    "oq_rhs_eval_device": (_int, [_vp, _d, _vp, _vp]),
    "oq_rhs_set_dilatancy": (_int, [_vp, ctypes.POINTER(DilaProps)]),
    "oq_rhs_eval_dila": (_int, [_vp, _d] + [_pd] * 8),
    "oq_rhs_info": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "oq_rhs_get_k": (_int, [_vp, _pd]),
    "oq_rhs_time": (_int, [_vp, _int, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "oq_rhs_destroy": (_int, [_vp]),
    "oq_stepper_create": (_int, [_vp, _pd, _d, _d, ctypes.POINTER(SolverOpts), ctypes.POINTER(_vp)]),
    "oq_stepper_create_vcabm5": (_int, [_vp, _pd, _d, _d, ctypes.POINTER(SolverOpts),
                                        ctypes.POINTER(_vp)]),
    "oq_stepper_run": (_int, [_vp, _i64, ctypes.POINTER(_int)]),
    "oq_stepper_state": (_int, [_vp, ctypes.POINTER(_d), ctypes.POINTER(_d), _pd,
                                ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "oq_stepper_destroy": (_int, [_vp]),
    "oq_stepper_set_persistent": (_int, [_vp, _int]),
    "oq_stepper_record": (_int, [_vp, _i64, _i64, _i64]),
    "oq_stepper_drain": (_int, [_vp, ctypes.POINTER(_i64), _pd, _pd, ctypes.POINTER(_i64), _pd, _pd]),
    "oq_stepper_record_du": (_int, [_vp, _i64, _i64, _i64, _int]),
    "oq_stepper_drain_du": (_int, [_vp, ctypes.POINTER(_i64), _pd, _pd, ctypes.POINTER(_i64), _pd, _pd,
                                   _pd]),
    "oq_ensemble_create": (_int, [_int, ctypes.POINTER(FaultMesh), ctypes.POINTER(Hex8Mesh), _d, _d,
                                  _int, _i64, _d, ctypes.POINTER(Quadrature), _i64,
                                  ctypes.POINTER(RSProps), ctypes.POINTER(ViscProps), _int,
                                  ctypes.POINTER(_vp)]),
    "oq_ensemble_eval": (_int, [_vp, _pd, _pd]),
    "oq_ensemble_start": (_int, [_vp, _pd, _d, _d, ctypes.POINTER(SolverOpts)]),
    "oq_ensemble_start_vcabm5": (_int, [_vp, _pd, _d, _d, ctypes.POINTER(SolverOpts)]),
    "oq_ensemble_get_ctrl": (_int, [_vp, ctypes.POINTER(LockstepCtrl)]),
    "oq_ensemble_resume": (_int, [_vp, _pd, ctypes.POINTER(LockstepCtrl), _d, ctypes.POINTER(SolverOpts)]),
    "oq_ensemble_run": (_int, [_vp, _i64, ctypes.POINTER(_i64)]),
    "oq_ensemble_state": (_int, [_vp, _pd, _pd, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                 ctypes.POINTER(_int)]),
    "oq_ensemble_record": (_int, [_vp, _i64]),
    "oq_ensemble_drain": (_int, [_vp, ctypes.POINTER(_i64), _pd, _pd]),
    "oq_ensemble_time": (_int, [_vp, _int, ctypes.POINTER(_d)]),
    "oq_ensemble_destroy": (_int, [_vp]),
    "oq_group_create": (_int, [_int, _i64, ctypes.POINTER(FaultMesh), _vp, _vp, _d, _d, _int, _i64,
                               _d, ctypes.POINTER(Quadrature), _i64, ctypes.POINTER(RSProps), _int,
                               ctypes.POINTER(_vp)]),
    "oq_group_info": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "oq_group_autotune": (_int, [_vp, _int, ctypes.POINTER(_int), _pd, _pd]),
    "oq_group_set_shapes": (_int, [_vp, ctypes.POINTER(_int)]),
    "oq_group_eval": (_int, [_vp, _pd, _pd]),
    "oq_group_time": (_int, [_vp, _int, ctypes.POINTER(_d)]),
    "oq_group_start": (_int, [_vp, _pd, _d, _d, ctypes.POINTER(SolverOpts)]),
    "oq_group_start_vcabm5": (_int, [_vp, _pd, _d, _d, ctypes.POINTER(SolverOpts)]),
    "oq_group_get_ctrl": (_int, [_vp, ctypes.POINTER(LockstepCtrl)]),
    "oq_group_resume": (_int, [_vp, _pd, ctypes.POINTER(LockstepCtrl), _d, ctypes.POINTER(SolverOpts)]),
    "oq_group_run": (_int, [_vp, _i64, ctypes.POINTER(_i64)]),
    "oq_group_state": (_int, [_vp, _pd, _pd, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                              ctypes.POINTER(_int)]),
    "oq_group_destroy": (_int, [_vp]),
}


def load():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise OetqfError("liboetqf_amd.so is not built (run __graft_entry__.build() or "
                             "make -C oetqf.jl_amd); there is no CPU fallback")
        lib = ctypes.CDLL(LIBPATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check_drain(rc):
    """Status of a recorder drain: an overflow still hands out the newest
    records (and restarts the recorder), so it becomes a RuntimeWarning
    naming how many were lost; any other error raises."""
    if rc == OQ_ERR_STATE:
        msg = load().oq_last_error().decode("utf-8", "replace")
        if "overflow" in msg:
            warnings.warn(msg, RuntimeWarning, stacklevel=3)
            return
    check(rc)


# Objects owning device handles.  They are closed by an exit hook, before the
# interpreter and the HIP runtime shut down, so that no stream, graph or
# buffer is released from a finalizer during teardown.
_live = weakref.WeakSet()


def register(obj):
    _live.add(obj)
    return obj


@atexit.register
def _close_live_handles():
    # steppers borrow their RHS handle: close them first
    for obj in sorted(list(_live), key=lambda o: getattr(o, "_close_rank", 1)):
        try:
            obj.close()
        except Exception:
            pass


def check(rc):
    if rc != OQ_OK:
        msg = load().oq_last_error().decode("utf-8", "replace")
        if rc == OQ_ERR_ARG:
            raise ValueError(msg)
        raise OetqfError("oetqf_amd error %d: %s" % (rc, msg))


def ptr(a):
    return a.ctypes.data_as(_pd)


def device_count():
    n = _int(0)
    check(load().oq_device_count(ctypes.byref(n)))
    return n.value
