"""CPU oracle (TEST INFRASTRUCTURE): ctypes bindings to oracle/lib/liboetqf_oracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, as the checker.  It restates the reference's algorithms
(src/BEM/mesh.jl, src/BEM/GF.jl, src/BEM/equation.jl) in FP64 C; the
docstrings cite the reference lines each wrapper follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# OQ_ORACLE_LIB selects another build of the same sources (the sanitizer
# build, lib/liboetqf_oracle_san.so)
_LIBPATH = os.environ.get("OQ_ORACLE_LIB") or os.path.join(_HERE, "lib", "liboetqf_oracle.so")
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_pd = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile the oracle library (gcc, plain C)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        _lib = ctypes.CDLL(_LIBPATH)
        _lib.oq_dc3d.restype = ctypes.c_int
        _lib.oq_dc3d.argtypes = [_d] * 13 + [_pd]
        _lib.oq_sind.restype = _d
        _lib.oq_sind.argtypes = [_d]
        _lib.oq_cosd.restype = _d
        _lib.oq_cosd.argtypes = [_d]
        _lib.oq_range_len.restype = _i64
        _lib.oq_range_len.argtypes = [_d, _d, _d]
        _lib.oq_stress_vol_hex8.restype = None
        _lib.oq_stress_vol_hex8.argtypes = [_d] * 9 + [_pd, _d, _d, ctypes.c_int, _pd]
        _lib.oq_tsit5_tableau.restype = _pd
    return _lib


def _p(a):
    return a.ctypes.data_as(_pd)


class FaultMeshC(ctypes.Structure):
    _fields_ = [("x", _pd), ("ax", _pd), ("y", _pd), ("z", _pd), ("axi", _pd),
                ("nx", _i64), ("nxi", _i64), ("dx", _d), ("dep", _d), ("dip", _d)]


class Hex8MeshC(ctypes.Structure):
    _fields_ = [("cx", _pd), ("cy", _pd), ("cz", _pd), ("qx", _pd), ("qy", _pd),
                ("qz", _pd), ("dx", _pd), ("dy", _pd), ("dz", _pd), ("ne", _i64)]


class RSC(ctypes.Structure):
    _fields_ = [("a", _pd), ("b", _pd), ("L", _pd), ("sigma", _pd),
                ("eta", _d), ("vpl", _d), ("f0", _d), ("v0", _d)]


class ViscC(ctypes.Structure):
    _fields_ = [("gamma", _pd), ("n", _pd), ("deps0", _d * 6)]


class OdeOpts(ctypes.Structure):
    _fields_ = [("reltol", _d), ("abstol", _d), ("dtmax", _d), ("dtmin", _d),
                ("dt0", _d), ("maxiters", _i64)]


class OdeStats(ctypes.Structure):
    _fields_ = [("naccept", _i64), ("nreject", _i64), ("nf", _i64), ("t", _d),
                ("dt", _d), ("retcode", ctypes.c_int)]


def sind(x):
    return lib().oq_sind(float(x))


def cosd(x):
    return lib().oq_cosd(float(x))


def dc3d(x, y, z, alpha, depth, dip, al1, al2, aw1, aw2, d1, d2, d3):
    """Okada (1992) DC3D; returns (u[12], status).  Argument order as the
    reference's call sites (GF.jl:49-53)."""
    u = np.zeros(12)
    ret = lib().oq_dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, d1, d2, d3, _p(u))
    return u, ret


def stress_vol_hex8(x, y, z, qx, qy, qz, T, L, W, eps, mu, nu, nquad=6):
    """Stress from a uniform eigenstrain in a cuboid (GF.jl:215-221)."""
    e = np.ascontiguousarray(eps, dtype=np.float64)
    s = np.zeros(6)
    lib().oq_stress_vol_hex8(x, y, z, qx, qy, qz, T, L, W, _p(e), mu, nu, int(nquad), _p(s))
    return s


def range_len(start, stop, step):
    return int(lib().oq_range_len(start, stop, step))


class _Keep:
    """Holds numpy arrays referenced by a ctypes struct."""

    def __init__(self):
        self.refs = []

    def arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.refs.append(a)
        return a


def fault_struct(m, keep):
    ax = keep.arr(np.asarray(m.ax, dtype=np.float64).reshape(-1))
    axi = keep.arr(np.asarray(m.axi, dtype=np.float64).reshape(-1))
    return FaultMeshC(_p(keep.arr(m.x)), _p(ax), _p(keep.arr(m.y)), _p(keep.arr(m.z)),
                      _p(axi), m.nx, m.nxi, m.dx, m.dep, m.dip)


def hex8_struct(m, keep):
    return Hex8MeshC(*[_p(keep.arr(getattr(m, k))) for k in
                       ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")], len(m.cx))


def mesh_rect_okada(x, xi, dx, dxi, dip):
    """gen_mesh(Val(:RectOkada), x, ξ, Δx, Δξ, dip) (mesh.jl:39-56) -> dict."""
    L = lib()
    nxi = L.oq_range_len(0.0, -xi + dxi, -dxi)
    nx = L.oq_range_len(-x / 2 + dx / 2, x / 2 - dx / 2, dx)
    xs = np.zeros(nx); ax = np.zeros(2 * nx)
    L.oq_mesh_strike.argtypes = [_d, _d, _i64, _pd, _pd]
    L.oq_mesh_strike(x, dx, nx, _p(xs), _p(ax))
    xis = np.zeros(nxi); axi = np.zeros(2 * nxi); y = np.zeros(nxi); z = np.zeros(nxi)
    L.oq_mesh_downdip.argtypes = [_d, _d, _d, _i64, _pd, _pd, _pd, _pd]
    L.oq_mesh_downdip(xi, dxi, dip, nxi, _p(xis), _p(axi), _p(y), _p(z))
    return dict(x=xs, dx=dx, nx=nx, ax=ax.reshape(nx, 2), xi=xis, dxi=dxi, nxi=nxi,
                axi=axi.reshape(nxi, 2), y=y, z=z, dep=0.0, dip=dip)


def mesh_hex8(llx, lly, llz, dx, dy, dz, nx, ny, nz, rfzh=None):
    """Structured gen_gmsh_mesh + gen_mesh(:BEMHex8Mesh) (mesh.jl:95-186) -> dict."""
    ne = nx * ny * nz
    out = {k: np.zeros(ne) for k in ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")}
    rf = np.ascontiguousarray(np.ones(nz) if rfzh is None else rfzh, dtype=np.float64)
    L = lib()
    L.oq_mesh_hex8.argtypes = [_d] * 6 + [_i64] * 3 + [_pd] * 10
    L.oq_mesh_hex8(llx, lly, llz, dx, dy, dz, nx, ny, nz, _p(rf),
                   *[_p(out[k]) for k in ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")])
    return out


def gf_ff(mesh, lam, mu, ftype=0, nrept=2, buffer_ratio=0.0):
    """Toeplitz K_ff st (nx, nξ, nξ), Julia order (GF.jl:31-58); returned
    as a numpy array indexed st[i, j, l]."""
    keep = _Keep()
    m = fault_struct(mesh, keep)
    st = np.zeros(mesh.nx * mesh.nxi * mesh.nxi)
    L = lib()
    L.oq_gf_ff.argtypes = [ctypes.POINTER(FaultMeshC), _d, _d, ctypes.c_int, _i64, _d, _pd]
    L.oq_gf_ff(ctypes.byref(m), lam, mu, ftype, nrept, buffer_ratio, _p(st))
    return st.reshape(mesh.nxi, mesh.nxi, mesh.nx).transpose(2, 1, 0)


def gf_mf(mf, ma, lam, mu, ftype=0, qc=(0.0, 0.0, 0.0), qw=(1.0,), nrept=2, buffer_ratio=0.0):
    """K_mf = gf₁₂ (6ne x nDisl) (GF.jl:123-174)."""
    keep = _Keep()
    a = fault_struct(mf, keep)
    b = hex8_struct(ma, keep)
    qc = keep.arr(np.asarray(qc, dtype=np.float64).reshape(-1))
    qw = keep.arr(np.asarray(qw, dtype=np.float64).reshape(-1))
    ne, nf = len(ma.cx), mf.nx * mf.nxi
    st = np.zeros(6 * ne * nf)
    L = lib()
    L.oq_gf_mf.argtypes = [ctypes.POINTER(FaultMeshC), ctypes.POINTER(Hex8MeshC), _d, _d,
                           ctypes.c_int, _pd, _pd, _i64, _i64, _d, _pd]
    L.oq_gf_mf(ctypes.byref(a), ctypes.byref(b), lam, mu, ftype, _p(qc), _p(qw), len(qw),
               nrept, buffer_ratio, _p(st))
    return st.reshape(nf, 6 * ne).T


def gf_fm(ma, mf, lam, mu, ftype=0, nquad=6):
    """K_fm = gf₂₁ (nDisl x 6ne) (GF.jl:194-227)."""
    keep = _Keep()
    a = fault_struct(mf, keep)
    b = hex8_struct(ma, keep)
    ne, nf = len(ma.cx), mf.nx * mf.nxi
    st = np.zeros(6 * ne * nf)
    L = lib()
    L.oq_gf_fm.argtypes = [ctypes.POINTER(Hex8MeshC), ctypes.POINTER(FaultMeshC), _d, _d,
                           ctypes.c_int, ctypes.c_int, _pd]
    L.oq_gf_fm(ctypes.byref(b), ctypes.byref(a), lam, mu, ftype, nquad, _p(st))
    return st.reshape(6 * ne, nf).T


def gf_mm(ma, lam, mu, qc=(0.0, 0.0, 0.0), qw=(1.0,), nquad=6):
    """K_mm = gf₂₂ (6ne x 6ne) (GF.jl:250-296)."""
    keep = _Keep()
    b = hex8_struct(ma, keep)
    qc = keep.arr(np.asarray(qc, dtype=np.float64).reshape(-1))
    qw = keep.arr(np.asarray(qw, dtype=np.float64).reshape(-1))
    ne = len(ma.cx)
    st = np.zeros(36 * ne * ne)
    L = lib()
    L.oq_gf_mm.argtypes = [ctypes.POINTER(Hex8MeshC), _d, _d, _pd, _pd, _i64, ctypes.c_int, _pd]
    L.oq_gf_mm(ctypes.byref(b), lam, mu, _p(qc), _p(qw), len(qw), nquad, _p(st))
    return st.reshape(6 * ne, 6 * ne).T


def gf_mm_rows(ma_src, ma_rcv, lam, mu, qc=(0.0, 0.0, 0.0), qw=(1.0,), nquad=6):
    """Rows of K_mm (GF.jl:250-296) for the receiver cells ma_rcv over the
    source cells ma_src: (6 nr) x (6 ns)."""
    keep = _Keep()
    a = hex8_struct(ma_src, keep)
    b = hex8_struct(ma_rcv, keep)
    qc = keep.arr(np.asarray(qc, dtype=np.float64).reshape(-1))
    qw = keep.arr(np.asarray(qw, dtype=np.float64).reshape(-1))
    ns, nr = len(ma_src.cx), len(ma_rcv.cx)
    st = np.zeros(36 * ns * nr)
    L = lib()
    L.oq_gf_mm_rows.argtypes = [ctypes.POINTER(Hex8MeshC), ctypes.POINTER(Hex8MeshC), _d, _d, _pd,
                                _pd, _i64, ctypes.c_int, _pd]
    L.oq_gf_mm_rows(ctypes.byref(a), ctypes.byref(b), lam, mu, _p(qc), _p(qw), len(qw), nquad,
                    _p(st))
    return st.reshape(6 * ns, 6 * nr).T


def _fortran_flat(a):
    return np.asarray(a, dtype=np.float64).reshape(-1, order="F")


def _rs(p, keep):
    """Friction parameters in Julia (column-major) order, like the state."""
    return RSC(_p(keep.arr(_fortran_flat(p.a))), _p(keep.arr(_fortran_flat(p.b))),
               _p(keep.arr(_fortran_flat(p.L))), _p(keep.arr(_fortran_flat(p.sigma))),
               p.eta, p.vpl, p.f0, p.v0)


def _visc(p, keep):
    c = ViscC(_p(keep.arr(p.gamma)), _p(keep.arr(np.asarray(p.n, dtype=np.float64))))
    for k in range(6):
        c.deps0[k] = float(p.deps0[k])
    return c


def update_fault(p, v, theta, dtau):
    """update_fault! (equation.jl:233-246)."""
    keep = _Keep()
    c = _rs(p, keep)
    n = v.size
    dv, dth, dd = np.zeros(n), np.zeros(n), np.zeros(n)
    L = lib()
    L.oq_update_fault.argtypes = [ctypes.POINTER(RSC), _i64, _pd, _pd, _pd, _pd, _pd, _pd]
    L.oq_update_fault(ctypes.byref(c), n, _p(keep.arr(v.reshape(-1, order="F"))),
                      _p(keep.arr(theta.reshape(-1, order="F"))),
                      _p(keep.arr(dtau.reshape(-1, order="F"))), _p(dv), _p(dth), _p(dd))
    return dv, dth, dd


class DilaC(ctypes.Structure):
    _fields_ = [("tp", _pd), ("eps", _pd), ("beta", _pd), ("p0", _pd)]


def _dila(d, keep):
    return DilaC(*[_p(keep.arr(_fortran_flat(getattr(d, k)))) for k in ("tp", "eps", "beta", "p0")])


def update_fault_dila(p, d, v, theta, pp, dtau):
    """update_fault_with_dilatancy! (equation.jl:248-276)."""
    keep = _Keep()
    c, cd = _rs(p, keep), _dila(d, keep)
    n = v.size
    out = [np.zeros(n) for _ in range(4)]
    L = lib()
    L.oq_update_fault_dila.argtypes = [ctypes.POINTER(RSC), ctypes.POINTER(DilaC), _i64] + [_pd] * 8
    L.oq_update_fault_dila(ctypes.byref(c), ctypes.byref(cd), n,
                           *[_p(keep.arr(x.reshape(-1, order="F"))) for x in (v, theta, pp, dtau)],
                           *[_p(o) for o in out])
    return tuple(out)  # dv, dtheta, ddelta, dp


def rhs_fault_dila(pf, d, st, v, theta, pp):
    """ode with dilatancy (equation.jl:168-183)."""
    keep = _Keep()
    c, cd = _rs(pf, keep), _dila(d, keep)
    nx, nxi = v.shape
    out = [np.zeros(nx * nxi) for _ in range(4)]
    L = lib()
    L.oq_rhs_fault_dila.argtypes = [ctypes.POINTER(RSC), ctypes.POINTER(DilaC), _pd, _i64, _i64] + [_pd] * 7
    L.oq_rhs_fault_dila(ctypes.byref(c), ctypes.byref(cd), _p(keep.arr(_st_flat(st))), nx, nxi,
                        *[_p(keep.arr(x.reshape(-1, order="F"))) for x in (v, theta, pp)],
                        *[_p(o) for o in out])
    sh = (nx, nxi)
    return tuple(o.reshape(sh, order="F") for o in out)  # dv, dtheta, ddelta, dp


def update_strain_rate(p, sigma):
    """update_strain_rate! (equation.jl:207-222); sigma (ne, 6)."""
    keep = _Keep()
    c = _visc(p, keep)
    ne = sigma.shape[0]
    de = np.zeros(6 * ne)
    L = lib()
    L.oq_update_strain_rate.argtypes = [ctypes.POINTER(ViscC), _i64, _pd, _pd]
    L.oq_update_strain_rate(ctypes.byref(c), ne, _p(keep.arr(sigma.reshape(-1, order="F"))), _p(de))
    return de.reshape(6, ne).T


def _st_flat(st):
    return np.ascontiguousarray(np.asarray(st).transpose(2, 1, 0)).reshape(-1)


def dtau_fft(st, s):
    """dτ_dt! (equation.jl:44-61) with the Toeplitz kernel st (nx,nξ,nξ);
    s (nx, nξ)."""
    nx, nxi = s.shape
    out = np.zeros(nx * nxi)
    L = lib()
    L.oq_dtau_fft.argtypes = [_pd, _i64, _i64, _pd, _pd]
    stf = _st_flat(st)
    sf = np.ascontiguousarray(s.reshape(-1, order="F"))
    L.oq_dtau_fft(_p(stf), nx, nxi, _p(sf), _p(out))
    return out.reshape(nx, nxi, order="F")


def dtau_toeplitz(st, s):
    nx, nxi = s.shape
    out = np.zeros(nx * nxi)
    L = lib()
    L.oq_dtau_toeplitz.argtypes = [_pd, _i64, _i64, _pd, _pd]
    stf = _st_flat(st)
    sf = np.ascontiguousarray(s.reshape(-1, order="F"))
    L.oq_dtau_toeplitz(_p(stf), nx, nxi, _p(sf), _p(out))
    return out.reshape(nx, nxi, order="F")


def rhs_coupled(pf, pa, st, gf12, gf21, gf22, v, theta, sigma):
    """ode(du,u,p,t) coupled (equation.jl:185-205).  Returns
    (dv, dθ, dϵ, dσ, dδ) with Julia shapes."""
    keep = _Keep()
    cf, ca = _rs(pf, keep), _visc(pa, keep)
    nx, nxi = v.shape
    ne = sigma.shape[0]
    nf = nx * nxi
    dv, dth, dd = np.zeros(nf), np.zeros(nf), np.zeros(nf)
    de, ds = np.zeros(6 * ne), np.zeros(6 * ne)
    L = lib()
    L.oq_rhs_coupled.argtypes = [ctypes.POINTER(RSC), ctypes.POINTER(ViscC)] + [_pd] * 4 + \
        [_i64] * 3 + [_pd] * 8
    L.oq_rhs_coupled(ctypes.byref(cf), ctypes.byref(ca), _p(keep.arr(_st_flat(st))),
                     _p(keep.arr(np.asarray(gf12).reshape(-1, order="F"))),
                     _p(keep.arr(np.asarray(gf21).reshape(-1, order="F"))),
                     _p(keep.arr(np.asarray(gf22).reshape(-1, order="F"))),
                     nx, nxi, ne, _p(keep.arr(v.reshape(-1, order="F"))),
                     _p(keep.arr(theta.reshape(-1, order="F"))),
                     _p(keep.arr(sigma.reshape(-1, order="F"))),
                     _p(dv), _p(dth), _p(de), _p(ds), _p(dd))
    sh = (nx, nxi)
    return (dv.reshape(sh, order="F"), dth.reshape(sh, order="F"),
            de.reshape(6, ne).T, ds.reshape(6, ne).T, dd.reshape(sh, order="F"))


def rhs_fault(pf, st, v, theta):
    """ode(du,u,p,t) fault only (equation.jl:156-166)."""
    keep = _Keep()
    cf = _rs(pf, keep)
    nx, nxi = v.shape
    nf = nx * nxi
    dv, dth, dd = np.zeros(nf), np.zeros(nf), np.zeros(nf)
    L = lib()
    L.oq_rhs_fault.argtypes = [ctypes.POINTER(RSC), _pd, _i64, _i64] + [_pd] * 5
    L.oq_rhs_fault(ctypes.byref(cf), _p(keep.arr(_st_flat(st))), nx, nxi,
                   _p(keep.arr(v.reshape(-1, order="F"))),
                   _p(keep.arr(theta.reshape(-1, order="F"))), _p(dv), _p(dth), _p(dd))
    sh = (nx, nxi)
    return dv.reshape(sh, order="F"), dth.reshape(sh, order="F"), dd.reshape(sh, order="F")


def rhs_dense(pf, pa, K, nf, ne, v, theta, sigma):
    """Coupled RHS with a dense receiver-major operator K (row-major N x N).
    v, theta flat (nf,); sigma flat component-major (6ne,)."""
    keep = _Keep()
    cf = _rs(pf, keep)
    ca = _visc(pa, keep) if ne > 0 else ViscC()
    dv, dth, dd = np.zeros(nf), np.zeros(nf), np.zeros(nf)
    de, ds = np.zeros(max(6 * ne, 1)), np.zeros(max(6 * ne, 1))
    L = lib()
    L.oq_rhs_dense.argtypes = [ctypes.POINTER(RSC), ctypes.POINTER(ViscC), _pd, _i64, _i64] + [_pd] * 8
    sig = keep.arr(sigma if ne > 0 else np.zeros(1))
    L.oq_rhs_dense(ctypes.byref(cf), ctypes.byref(ca), _p(keep.arr(K.reshape(-1))), nf, ne,
                   _p(keep.arr(v)), _p(keep.arr(theta)), _p(sig), _p(dv), _p(dth), _p(de),
                   _p(ds), _p(dd))
    return dv, dth, de[:6 * ne], ds[:6 * ne], dd


RHSFN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _d, _pd, _pd)


def tsit5(f, u0, t0, t1, reltol=1e-6, abstol=1e-8, dtmax=np.inf, dtmin=0.0, dt0=0.0,
          maxiters=10**9):
    """Tsit5 with OrdinaryDiffEq's controller (oracle/tsit5.c).  f(t, u) -> du
    on flat numpy arrays.  Returns (u, stats)."""
    u = np.array(u0, dtype=np.float64, copy=True)
    n = u.size

    def cb(_ctx, t, up, dup):
        uu = np.ctypeslib.as_array(up, shape=(n,))
        du = np.ctypeslib.as_array(dup, shape=(n,))
        du[:] = f(t, uu)

    fn = RHSFN(cb)
    o = OdeOpts(reltol, abstol, dtmax, dtmin, dt0, maxiters)
    st = OdeStats()
    L = lib()
    L.oq_tsit5.argtypes = [RHSFN, ctypes.c_void_p, _i64, _pd, _d, _d, ctypes.POINTER(OdeOpts),
                           ctypes.POINTER(OdeStats)]
    L.oq_tsit5(fn, None, n, _p(u), t0, t1, ctypes.byref(o), ctypes.byref(st))
    return u, st


def vcabm5(f, u0, t0, t1, reltol=1e-6, abstol=1e-8, dtmax=np.inf, dtmin=0.0, dt0=0.0,
           maxiters=10**9):
    """VCABM5 (oracle/vcabm.c).  f(t, u) -> du on flat numpy arrays.
    Returns (u, stats)."""
    u = np.array(u0, dtype=np.float64, copy=True)
    n = u.size

    def cb(_ctx, t, up, dup):
        uu = np.ctypeslib.as_array(up, shape=(n,))
        du = np.ctypeslib.as_array(dup, shape=(n,))
        du[:] = f(t, uu)

    fn = RHSFN(cb)
    o = OdeOpts(reltol, abstol, dtmax, dtmin, dt0, maxiters)
    st = OdeStats()
    L = lib()
    L.oq_vcabm5.argtypes = [RHSFN, ctypes.c_void_p, _i64, _pd, _d, _d, ctypes.POINTER(OdeOpts),
                            ctypes.POINTER(OdeStats)]
    L.oq_vcabm5(fn, None, n, _p(u), t0, t1, ctypes.byref(o), ctypes.byref(st))
    return u, st


def vcabm_coeffs(k, h, hist):
    """(g[0..k], beta[0..k-1]) of the VCABM step (oracle/vcabm.c)."""
    hs = np.zeros(6)
    hs[:len(hist)] = hist
    g, beta = np.zeros(6), np.zeros(6)
    L = lib()
    L.oq_vcabm_coeffs.argtypes = [ctypes.c_int, _d, _pd, _pd, _pd]
    L.oq_vcabm_coeffs(int(k), float(h), _p(hs), _p(g), _p(beta))
    return g[:k + 1].copy(), beta[:k].copy()


def tsit5_tableau():
    p = lib().oq_tsit5_tableau()
    a = np.ctypeslib.as_array(p, shape=(70,)).copy()
    return a[:7], a[7:56].reshape(7, 7), a[56:63], a[63:70]


def cycle_dense(pf, pa, K, nf, ne, u0, t0, t1, reltol=1e-6, abstol=1e-8, dtmax=np.inf, dtmin=0.0,
                dt0=0.0, maxiters=10**9, cap=2_000_000, alg="tsit5"):
    """A whole cycle in C (oq_cycle_dense_alg): Tsit5 or VCABM5 over the
    dense RHS; returns (u, stats, step_t, step_vmax) for every accepted
    step."""
    keep = _Keep()
    cf = _rs(pf, keep)
    ca = _visc(pa, keep) if ne > 0 else ViscC()
    u = np.array(u0, dtype=np.float64, copy=True)
    st_t, st_v = np.zeros(cap), np.zeros(cap)
    nrec = ctypes.c_int64()
    o = OdeOpts(reltol, abstol, dtmax, dtmin, dt0, maxiters)
    st = OdeStats()
    L = lib()
    L.oq_cycle_dense_alg.argtypes = [ctypes.POINTER(RSC), ctypes.POINTER(ViscC), _pd, _i64, _i64,
                                     _pd, _d, _d, ctypes.POINTER(OdeOpts), ctypes.POINTER(OdeStats),
                                     _pd, _pd, _i64, ctypes.POINTER(_i64), ctypes.c_int]
    Kc = keep.arr(np.asarray(K).reshape(-1))
    L.oq_cycle_dense_alg(ctypes.byref(cf), ctypes.byref(ca), _p(Kc), nf, ne, _p(u), t0, t1,
                         ctypes.byref(o), ctypes.byref(st), _p(st_t), _p(st_v), cap,
                         ctypes.byref(nrec), 1 if alg == "vcabm5" else 0)
    k = min(nrec.value, cap)
    return u, st, st_t[:k].copy(), st_v[:k].copy()
