"""ODE assembly, RHS and solve (src/BEM/equation.jl, the OrdinaryDiffEq
calls of src/io.jl) on the GPU.

`assemble` binds the Green's functions and properties to a device-resident
RHS handle (the analogue of the reference's ODEProblem with preallocated
scratch, equation.jl:81-89, :141-154).  `prob.f(du, u, p, t)` evaluates the
RHS in the gfx950 kernels (equation.jl:156-166, :185-205); `solve(prob,
Tsit5(); ...)` integrates on the device with the Tsit5 stepper.
"""
import ctypes
import weakref

import numpy as np

from . import _lib
from .greens import StrikeSlip, _Keep, _ftype, fault_struct, get_quadrature, hex8_struct, quad_struct
from .property import DieterichStateLaw


class ArrayPartition:
    """The reference's state container (RecursiveArrayTools.ArrayPartition):
    a tuple of Fortran-ordered float64 arrays, packed in order."""

    def __init__(self, *xs):
        self.x = tuple(np.asfortranarray(np.asarray(a, dtype=np.float64)) for a in xs)

    def similar(self):
        return ArrayPartition(*[np.zeros_like(a, order="F") for a in self.x])

    def pack(self):
        return np.concatenate([a.reshape(-1, order="F") for a in self.x])

    def unpack(self, flat):
        off = 0
        for a in self.x:
            n = a.size
            a.reshape(-1, order="F")[:] = flat[off:off + n]
            off += n
        return self

    @property
    def size(self):
        return sum(a.size for a in self.x)


def similar(u):
    return u.similar()


def _rs_struct(p, keep):
    return _lib.RSProps(_lib.ptr(keep.arr(np.asarray(p.a).reshape(-1, order="F"))),
                        _lib.ptr(keep.arr(np.asarray(p.b).reshape(-1, order="F"))),
                        _lib.ptr(keep.arr(np.asarray(p.L).reshape(-1, order="F"))),
                        _lib.ptr(keep.arr(np.asarray(p.σ).reshape(-1, order="F"))),
                        p.η, p.vpl, p.f0, p.v0)


def _visc_struct(pa, ne, keep):
    laws = pa.laws()
    g = np.concatenate([np.broadcast_to(l.γ, (ne,)) for l in laws])
    n = np.concatenate([np.broadcast_to(l.n, (ne,)) for l in laws])
    v = _lib.ViscProps(_lib.ptr(keep.arr(g)), _lib.ptr(keep.arr(n)), len(laws))
    for k in range(6):
        v.deps0[k] = float(pa.dϵ0[k])
    return v


def _dila_struct(d, keep):
    return _lib.DilaProps(*[_lib.ptr(keep.arr(np.asarray(getattr(d, k)).reshape(-1, order="F")))
                            for k in ("tp", "eps", "beta", "p0")])


class RHS:
    """Owns an oq_rhs handle; callable as f(du, u, p, t)."""

    def __init__(self, handle, nx, nxi, ne, keep, dila=None):
        self.h = handle
        self.nx, self.nxi, self.ne = nx, nxi, ne
        self._keep = keep
        self._steppers = weakref.WeakSet()  # they borrow this handle
        _lib.register(self)
        self.dila = dila is not None
        if self.dila:
            ds = _dila_struct(dila, keep)
            keep.append(ds)
            _lib.check(_lib.load().oq_rhs_set_dilatancy(handle, ctypes.byref(ds)))

    def __call__(self, du, u, p=None, t=0.0):
        L = _lib.load()
        ux, dux = u.x, du.x
        for a in list(ux) + list(dux):
            if not a.flags.f_contiguous or a.dtype != np.float64:
                raise TypeError("ArrayPartition components must be Fortran-ordered float64")
        P = _lib.ptr
        if self.dila:
            v, th, dl, pp = ux
            dv, dth, ddl, dpp = dux
            _lib.check(L.oq_rhs_eval_dila(self.h, float(t), P(v), P(th), P(dl), P(pp),
                                          P(dv), P(dth), P(ddl), P(dpp)))
        elif self.ne == 0:
            v, th, dl = ux
            dv, dth, ddl = dux
            _lib.check(L.oq_rhs_eval(self.h, float(t), P(v), P(th), None, None, P(dl),
                                     P(dv), P(dth), None, None, P(ddl)))
        else:
            v, th, ep, sg, dl = ux
            dv, dth, dep, dsg, ddl = dux
            _lib.check(L.oq_rhs_eval(self.h, float(t), P(v), P(th), P(ep), P(sg), P(dl),
                                     P(dv), P(dth), P(dep), P(dsg), P(ddl)))
        return du

    def info(self):
        ns, n, kb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.load().oq_rhs_info(self.h, ctypes.byref(ns), ctypes.byref(n),
                                           ctypes.byref(kb)))
        return ns.value, n.value, kb.value

    def dense_operator(self):
        _, n, _ = self.info()
        k = np.zeros(n * n)
        _lib.check(_lib.load().oq_rhs_get_k(self.h, _lib.ptr(k)))
        return k.reshape(n, n)

    def time(self, iters=10):
        a, b = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.load().oq_rhs_time(self.h, int(iters), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        for it in list(self._steppers):
            it.close()
        if self.h:
            _lib.load().oq_rhs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ODEProblem:
    """ODEProblem{true}(f, u0, tspan, p) analogue."""

    def __init__(self, f, u0, tspan, p):
        self.f, self.u0, self.tspan, self.p = f, u0, (float(tspan[0]), float(tspan[1])), p


_KDTYPES = {"fp64": _lib.OQ_K_FP64, "fp32": _lib.OQ_K_FP32, "bf16": _lib.OQ_K_BF16,
            "toeplitz": _lib.OQ_K_TOEPLITZ, "fft": _lib.OQ_K_FFT}
# assemble_from_mesh only: "fft_grid" = "fft" with K_mm applied as 2D
# convolutions over a mantle whose cells form a uniform horizontal grid
_KDTYPES_MESH = dict(_KDTYPES, fft_grid=_lib.OQ_K_FFT_GRID)


def assemble(*args, se=DieterichStateLaw(), kdtype="fp64", device=0, **kwargs):
    """assemble(gf, p, u0, tspan) (equation.jl:81-89),
    assemble(gf, p, dila, u0, tspan) (equation.jl:108-117, u0 = (v, θ, δ, 𝓅)) or
    assemble(gf₁₁, gf₁₂, gf₂₁, gf₂₂, pf, pa, u0, tspan) (equation.jl:141-154).
    gf/gf₁₁ may be the real Toeplitz kernel or its DFT (the reference
    builder's default, fourier=true)."""
    if not isinstance(se, DieterichStateLaw):
        raise NotImplementedError("only the aging law (DieterichStateLaw) is provided")
    if kdtype not in _KDTYPES:
        raise ValueError("kdtype must be one of %s" % sorted(_KDTYPES))
    kdt = _KDTYPES[kdtype]
    keep = _Keep()
    L = _lib.load()
    h = ctypes.c_void_p()
    dila = None
    if len(args) == 4:
        gf, pf, u0, tspan = args
        pa, gf12, gf21, gf22, ne = None, None, None, None, 0
    elif len(args) == 5:
        gf, pf, dila, u0, tspan = args
        pa, gf12, gf21, gf22, ne = None, None, None, None, 0
    elif len(args) == 8:
        gf, gf12, gf21, gf22, pf, pa, u0, tspan = args
        ne = u0.x[2].shape[0]
    else:
        raise TypeError("assemble takes (gf, p, u0, tspan), (gf, p, dila, u0, tspan) or "
                        "(gf₁₁, gf₁₂, gf₂₁, gf₂₂, pf, pa, u0, tspan)")
    nx, nxi = u0.x[0].shape
    is_dft = np.iscomplexobj(gf)
    if is_dft:
        g = np.asarray(gf, dtype=np.complex128).reshape(-1, order="F")
        g11 = keep.arr(np.stack([g.real, g.imag], axis=1).reshape(-1))
    else:
        g11 = keep.arr(np.asarray(gf).reshape(-1, order="F"))
    rs = _rs_struct(pf, keep)
    if ne > 0:
        vs = _visc_struct(pa, ne, keep)
        p12 = _lib.ptr(keep.arr(np.asarray(gf12).reshape(-1, order="F")))
        p21 = _lib.ptr(keep.arr(np.asarray(gf21).reshape(-1, order="F")))
        p22 = _lib.ptr(keep.arr(np.asarray(gf22).reshape(-1, order="F")))
        _lib.check(L.oq_rhs_create(device, nx, nxi, ne, ctypes.byref(rs), ctypes.byref(vs),
                                   _lib.ptr(g11), int(is_dft), p12, p21, p22, kdt,
                                   ctypes.byref(h)))
    else:
        _lib.check(L.oq_rhs_create(device, nx, nxi, 0, ctypes.byref(rs), None, _lib.ptr(g11),
                                   int(is_dft), None, None, None, kdt, ctypes.byref(h)))
    f = RHS(h, nx, nxi, ne, keep, dila=dila)
    return ODEProblem(f, u0, tspan, (pf, dila if dila is not None else pa, f, se))


def assemble_from_mesh(mf, ma, lam, mu, pf, pa, u0, tspan, ftype=StrikeSlip(), nrept=2,
                       buffer_ratio=0.0, qtype="Gauss1", kdtype="fp64", device=0, dila=None):
    """The four stress_greens_function calls followed by assemble, with the
    Green's functions built straight into the device operator (no host copy
    of K).  ma = None gives the fault-only problem.  kdtype "fft_grid"
    (coupled, mantle cells on a uniform horizontal grid): K_ff by DFT and
    K_mm as 2D convolutions of its per-offset kernel; K_fm, K_mf dense."""
    if kdtype not in _KDTYPES_MESH:
        raise ValueError("kdtype must be one of %s" % sorted(_KDTYPES_MESH))
    keep = _Keep()
    L = _lib.load()
    h = ctypes.c_void_p()
    a = fault_struct(mf, keep)
    rs = _rs_struct(pf, keep)
    if ma is not None:
        b = hex8_struct(ma, keep)
        q = quad_struct(get_quadrature(qtype), keep)
        vs = _visc_struct(pa, ma.ne, keep)
        _lib.check(L.oq_rhs_create_from_mesh(device, ctypes.byref(a), ctypes.byref(b), lam, mu,
                                             _ftype(ftype), nrept, buffer_ratio, ctypes.byref(q),
                                             ctypes.byref(rs), ctypes.byref(vs),
                                             _KDTYPES_MESH[kdtype], ctypes.byref(h)))
        ne = ma.ne
    else:
        _lib.check(L.oq_rhs_create_from_mesh(device, ctypes.byref(a), None, lam, mu,
                                             _ftype(ftype), nrept, buffer_ratio, None,
                                             ctypes.byref(rs), None, _KDTYPES_MESH[kdtype],
                                             ctypes.byref(h)))
        ne = 0
    f = RHS(h, mf.nx, mf.nxi, ne, keep, dila=dila)
    return ODEProblem(f, u0, tspan, (pf, dila if dila is not None else pa, f, DieterichStateLaw()))


class Tsit5:
    """Tsitouras 5(4) (the algorithm object passed to solve)."""


class VCABM5:
    """Variable-coefficient Adams-Bashforth-Moulton of order 5, PECE (the
    solver of the reference's example, examples/otf-with-mantle.jl:160)."""


class ODESolution:
    def __init__(self, t, u, retcode, naccept, nreject):
        self.t, self.u, self.retcode = t, u, retcode
        self.naccept, self.nreject = naccept, nreject


class Integrator:
    """Device-resident Tsit5 integration of an assembled problem."""

    _close_rank = 0  # closed before the RHS handle it borrows

    def __init__(self, prob, reltol=1e-6, abstol=1e-8, dt=0.0, dtmax=0.0, dtmin=0.0,
                 maxiters=10**9, alg=None):
        self.prob = prob
        self.u0 = prob.u0
        self._u0flat = np.ascontiguousarray(prob.u0.pack())
        o = _lib.SolverOpts(reltol, abstol, dt, dtmax, dtmin, int(maxiters))
        self.s = ctypes.c_void_p()
        if alg is not None and not isinstance(alg, (Tsit5, VCABM5)):
            raise NotImplementedError("the device steppers are Tsit5 and VCABM5")
        create = (_lib.load().oq_stepper_create_vcabm5 if isinstance(alg, VCABM5)
                  else _lib.load().oq_stepper_create)
        if not prob.f.h:
            raise ValueError("the problem's RHS handle is closed")
        _lib.check(create(prob.f.h, _lib.ptr(self._u0flat), prob.tspan[0], prob.tspan[1],
                          ctypes.byref(o), ctypes.byref(self.s)))
        prob.f._steppers.add(self)
        _lib.register(self)

    def step(self, nsteps):
        rc = ctypes.c_int()
        _lib.check(_lib.load().oq_stepper_run(self.s, int(nsteps), ctypes.byref(rc)))
        return _lib.RETCODES[rc.value]

    def state(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        na, nr = ctypes.c_int64(), ctypes.c_int64()
        flat = np.zeros(self._u0flat.size)
        _lib.check(_lib.load().oq_stepper_state(self.s, ctypes.byref(t), ctypes.byref(dt),
                                                _lib.ptr(flat), ctypes.byref(na),
                                                ctypes.byref(nr)))
        return t.value, dt.value, self.u0.similar().unpack(flat), na.value, nr.value

    def set_persistent(self, on=True):
        """Runs whole chunks of steps in one cooperative launch (fault-only
        problems whose operator rows fit in LDS); ValueError otherwise."""
        _lib.check(_lib.load().oq_stepper_set_persistent(self.s, 1 if on else 0))

    def record(self, stride=1, snap_capacity=0, step_capacity=4096, derivative=False):
        """Enables the device recorder: per accepted step (t, max|v|), and a
        state snapshot every `stride` accepted steps; with derivative=True
        each snapshot also keeps f(u) at the saved state."""
        _lib.check(_lib.load().oq_stepper_record_du(self.s, int(stride), int(snap_capacity),
                                                    int(step_capacity), 1 if derivative else 0))
        self._caps = (int(snap_capacity), int(step_capacity))
        self._du = bool(derivative)

    def drain(self):
        """Returns (step_t, step_vmax, snap_t, snaps (k, nstate)) recorded
        since the last drain, plus the derivative snapshots (k, nstate) when
        recorded with derivative=True.  A ring overrun hands out the newest
        records with a RuntimeWarning naming how many were lost."""
        sc, kc = self._caps
        st_t, st_v = np.zeros(max(kc, 1)), np.zeros(max(kc, 1))
        n = self._u0flat.size
        sn_t, sn_u = np.zeros(max(sc, 1)), np.zeros(max(sc, 1) * n)
        sn_du = np.zeros(max(sc, 1) * n) if self._du else None
        ns, nn = ctypes.c_int64(), ctypes.c_int64()
        _lib.check_drain(_lib.load().oq_stepper_drain_du(
            self.s, ctypes.byref(ns), _lib.ptr(st_t), _lib.ptr(st_v), ctypes.byref(nn),
            _lib.ptr(sn_t), _lib.ptr(sn_u), _lib.ptr(sn_du) if sn_du is not None else None))
        k, j = ns.value, nn.value
        out = (st_t[:k].copy(), st_v[:k].copy(), sn_t[:j].copy(), sn_u[:j * n].reshape(j, n).copy())
        if self._du:
            out += (sn_du[:j * n].reshape(j, n).copy(),)
        return out

    def close(self):
        if self.s:
            _lib.load().oq_stepper_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(prob, alg=None, reltol=1e-6, abstol=1e-8, dt=0.0, dtmax=0.0, dtmin=0.0,
          maxiters=10**9, chunk=4096):
    """solve(prob, Tsit5() or VCABM5(); reltol, abstol, dt, dtmax, maxiters):
    integrates to tspan[2] on the device and returns the final state."""
    it = Integrator(prob, reltol, abstol, dt, dtmax, dtmin, maxiters, alg=alg)
    ret = "Running"
    while ret == "Running":
        ret = it.step(chunk)
    t, _, u, na, nr = it.state()
    it.close()
    return ODESolution(t, u, ret, na, nr)
