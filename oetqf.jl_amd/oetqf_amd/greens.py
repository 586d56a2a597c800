"""Stress Green's functions (src/BEM/GF.jl) on the GPU.

`stress_greens_function` dispatches on its argument types exactly as the
reference's four methods do (GF.jl:31, :123, :194, :250); each call runs the
corresponding gfx950 builder through the C ABI and returns a freshly
allocated array in Julia layout (numpy Fortran order).
"""
import ctypes

import numpy as np

from . import _lib
from .mesh import BEMHex8Mesh, RectOkadaMesh


class FaultType:
    code = None


class StrikeSlip(FaultType):
    """StrikeSlip (GF.jl:4)."""
    code = 0


class DipSlip(FaultType):
    """DipSlip (GF.jl:5)."""
    code = 1


def _ftype(ftype):
    if isinstance(ftype, type) and issubclass(ftype, FaultType):
        return ftype.code
    if isinstance(ftype, FaultType):
        return ftype.code
    raise TypeError("ftype must be StrikeSlip() or DipSlip()")


class _Keep(list):
    def arr(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        self.append(a)
        return a


def fault_struct(m, keep):
    return _lib.FaultMesh(_lib.ptr(keep.arr(m.x)), _lib.ptr(keep.arr(m.ax)),
                          _lib.ptr(keep.arr(m.y)), _lib.ptr(keep.arr(m.z)),
                          _lib.ptr(keep.arr(m.axi)), m.nx, m.nxi, m.dx, m.dep, m.dip)


def hex8_struct(m, keep):
    return _lib.Hex8Mesh(*[_lib.ptr(keep.arr(getattr(m, k))) for k in
                           ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")], m.ne)


def quad_struct(q, keep):
    c, w = q
    c = keep.arr(c)
    w = keep.arr(w)
    assert c.size == 3 * w.size, "Wrong format of quadrature!"  # GF.jl:326
    return _lib.Quadrature(_lib.ptr(c), _lib.ptr(w), w.size)


def gmsh_quadrature(etype, qtype):
    """Hex (etype 5) integration points of gmsh's "GaussN" rules: the tensor
    Gauss-Legendre rule exact to degree N, i.e. ceil((N+1)/2) points per
    axis (gmsh >= 4.9, examples/otf-with-mantle.jl:64-66), in [-1,1]^3."""
    if etype != 5:
        raise NotImplementedError("only hex8 (etype 5) quadrature is provided")
    if not qtype.startswith("Gauss"):
        raise ValueError("unknown quadrature %r" % qtype)
    order = int(qtype[5:])
    npt = (order + 2) // 2
    x, w = np.polynomial.legendre.leggauss(npt)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    W = np.einsum("i,j,k->ijk", w, w, w)
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).ravel()
    return coords, W.ravel()


def get_quadrature(qtype):
    """get_quadrature (GF.jl:318-328): gmsh rule with weights normalised to
    sum 1, or a user tuple (coords, weights) checked for shape."""
    if isinstance(qtype, str):
        c, w = gmsh_quadrature(5, qtype)
        return c, w / w.sum()
    c, w = qtype
    c = np.asarray(c, dtype=np.float64).reshape(-1)
    w = np.asarray(w, dtype=np.float64).reshape(-1)
    assert c.size == 3 * w.size, "Wrong format of quadrature!"
    return c, w


def _gf_ff(mesh, lam, mu, ftype=StrikeSlip(), fourier=True, nrept=2, buffer_ratio=0.0):
    assert buffer_ratio >= 0, "Argument `buffer_ratio` must be ≥ 0."
    keep = _Keep()
    m = fault_struct(mesh, keep)
    L = _lib.load()
    if fourier:
        out = np.zeros(2 * mesh.nx * mesh.nxi * mesh.nxi)
        _lib.check(L.oq_gf_okada_ff_dft(ctypes.byref(m), lam, mu, _ftype(ftype), nrept,
                                        buffer_ratio, _lib.ptr(out)))
        c = out[0::2] + 1j * out[1::2]
        return c.reshape((mesh.nx, mesh.nxi, mesh.nxi), order="F")
    out = np.zeros(mesh.nx * mesh.nxi * mesh.nxi)
    _lib.check(L.oq_gf_okada_ff(ctypes.byref(m), lam, mu, _ftype(ftype), nrept, buffer_ratio,
                                _lib.ptr(out)))
    return out.reshape((mesh.nx, mesh.nxi, mesh.nxi), order="F")


def _gf_mf(mf, ma, lam, mu, ftype=StrikeSlip(), qtype="Gauss1", nrept=2, buffer_ratio=0.0):
    assert buffer_ratio >= 0, "Argument `buffer_ratio` must be ≥ 0."
    keep = _Keep()
    a, b = fault_struct(mf, keep), hex8_struct(ma, keep)
    q = quad_struct(get_quadrature(qtype), keep)
    nf = mf.nx * mf.nxi
    out = np.zeros(6 * ma.ne * nf)
    _lib.check(_lib.load().oq_gf_okada_mf(ctypes.byref(a), ctypes.byref(b), lam, mu,
                                          _ftype(ftype), ctypes.byref(q), nrept, buffer_ratio,
                                          _lib.ptr(out)))
    return out.reshape((6 * ma.ne, nf), order="F")


def _gf_fm(ma, mf, lam, mu, ftype=StrikeSlip()):
    keep = _Keep()
    a, b = fault_struct(mf, keep), hex8_struct(ma, keep)
    nf = mf.nx * mf.nxi
    out = np.zeros(6 * ma.ne * nf)
    _lib.check(_lib.load().oq_gf_hex8_fm(ctypes.byref(b), ctypes.byref(a), lam, mu, _ftype(ftype),
                                         _lib.ptr(out)))
    return out.reshape((nf, 6 * ma.ne), order="F")


def _gf_mm(ma, lam, mu, qtype="Gauss1", checkeigvals=True):
    keep = _Keep()
    b = hex8_struct(ma, keep)
    q = quad_struct(get_quadrature(qtype), keep)
    out = np.zeros(36 * ma.ne * ma.ne)
    _lib.check(_lib.load().oq_gf_hex8_mm(ctypes.byref(b), lam, mu, ctypes.byref(q),
                                         _lib.ptr(out)))
    st = out.reshape((6 * ma.ne, 6 * ma.ne), order="F")
    if checkeigvals:  # GF.jl:291-294, host-side diagnostic
        maxr = float(np.max(np.real(np.linalg.eigvals(st))))
        print("Maximum real part of eigval is: %.4f" % maxr)
    return st


def stress_greens_function(*args, **kwargs):
    """Dispatch of the four reference methods:
      (RectOkadaMesh, λ, μ; ftype, fourier=True, nrept=2, buffer_ratio=0)  GF.jl:31
      (RectOkadaMesh, BEMHex8Mesh, λ, μ; ftype, qtype, nrept, buffer_ratio) GF.jl:123
      (BEMHex8Mesh, RectOkadaMesh, λ, μ; ftype)                            GF.jl:194
      (BEMHex8Mesh, λ, μ; qtype, checkeigvals=True)                         GF.jl:250
    """
    kwargs.pop("fftw_flags", None)
    a0 = args[0]
    if isinstance(a0, RectOkadaMesh):
        if len(args) > 1 and isinstance(args[1], BEMHex8Mesh):
            return _gf_mf(*args, **kwargs)
        return _gf_ff(*args, **kwargs)
    if isinstance(a0, BEMHex8Mesh):
        if len(args) > 1 and isinstance(args[1], RectOkadaMesh):
            return _gf_fm(*args, **kwargs)
        return _gf_mm(*args, **kwargs)
    raise TypeError("no method matching stress_greens_function(%s)" % type(a0).__name__)
