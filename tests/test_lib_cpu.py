"""The C-ABI library without a GPU: it is built, loads, exports every symbol
include/oetqf_amd.h declares, and rejects bad arguments before touching a
device (the reference's @assert / ArgumentError behaviour)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "oetqf_amd.h")).read()
    # function declarations only: "<type> oq_name(" at the start of a line
    return sorted(set(re.findall(r"^(?:int|void|const char \*)\s*(oq_[a-z0-9_]+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol(oq):
    lib = oq.load_library()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    from oetqf_amd import _lib
    assert set(_lib.PROTOTYPES) == set(syms)


def _c_kind(p):
    p = p.strip()
    if "*" in p or "[" in p:
        return "ptr"
    for kind in ("double", "int64_t", "int"):
        if re.match(r"(const\s+)?%s\b" % kind, p):
            return kind
    return "unknown: " + p


def _ctypes_kind(t):
    if t in (ctypes.c_void_p, ctypes.c_char_p) or (isinstance(t, type) and issubclass(t, ctypes._Pointer)):
        return "ptr"
    return {ctypes.c_double: "double", ctypes.c_int64: "int64_t", ctypes.c_int: "int"}.get(t, repr(t))


def test_bindings_match_declared_signatures(oq):
    """Every ctypes prototype has the header's return type and, parameter
    by parameter, its kind (pointer, double, int64_t, int)."""
    from oetqf_amd import _lib
    src = open(os.path.join(ROOT, "include", "oetqf_amd.h")).read()
    decls = re.findall(r"^(int|void|const char \*)\s*(oq_[a-z0-9_]+)\s*\(([^;]*?)\)\s*;", src, re.M | re.S)
    assert len(decls) == len(declared_symbols())
    rest = {"int": ctypes.c_int, "void": None, "const char *": ctypes.c_char_p}
    for ret, name, params in decls:
        params = " ".join(params.split())
        want = [] if params in ("", "void") else [_c_kind(p) for p in params.split(",")]
        res, args = _lib.PROTOTYPES[name]
        assert res is rest[ret], name
        assert [_ctypes_kind(t) for t in args] == want, name


def test_version_and_error_string(oq):
    lib = oq.load_library()
    assert lib.oq_version() >= 1
    assert isinstance(lib.oq_last_error(), bytes)


def test_buffer_ratio_assert(oq):
    """@assert buffer_ratio ≥ 0 (GF.jl:36) -> ValueError from the C ABI."""
    m = oq.gen_mesh("RectOkada", 20e3, 10e3, 2.5e3, 2.5e3, 90.0)
    from oetqf_amd import _lib
    from oetqf_amd.greens import _Keep, fault_struct
    keep = _Keep()
    fs = fault_struct(m, keep)
    out = np.zeros(m.nx * m.nxi * m.nxi)
    rc = oq.load_library().oq_gf_okada_ff(ctypes.byref(fs), 3e10, 3e10, 0, 2, -1.0, _lib.ptr(out))
    assert rc == _lib.OQ_ERR_ARG
    assert "buffer_ratio" in oq.load_library().oq_last_error().decode()
    with pytest.raises(ValueError):
        _lib.check(rc)


def test_bad_ftype_and_null_mesh(oq):
    from oetqf_amd import _lib
    lib = oq.load_library()
    out = np.zeros(8)
    assert lib.oq_gf_okada_ff(None, 1.0, 1.0, 0, 2, 0.0, _lib.ptr(out)) == _lib.OQ_ERR_ARG
    m = oq.gen_mesh("RectOkada", 20e3, 10e3, 2.5e3, 2.5e3, 90.0)
    from oetqf_amd.greens import _Keep, fault_struct
    keep = _Keep()
    fs = fault_struct(m, keep)
    assert lib.oq_gf_okada_ff(ctypes.byref(fs), 1.0, 1.0, 7, 2, 0.0, _lib.ptr(out)) == _lib.OQ_ERR_ARG


def test_rhs_create_rejects_bad_properties(oq):
    """RateStateQuasiDynamicProperty invariants (property.jl:20-24) are checked
    by the C ABI before any device work."""
    from oetqf_amd import _lib
    lib = oq.load_library()
    a = np.ones(4)
    rs = _lib.RSProps(_lib.ptr(a), _lib.ptr(a), _lib.ptr(a), _lib.ptr(a), -1.0, 1e-9, 0.6, 1e-6)
    h = ctypes.c_void_p()
    st = np.zeros(8)
    rc = lib.oq_rhs_create(0, 2, 2, 0, ctypes.byref(rs), None, _lib.ptr(st), 0, None, None, None, 0,
                           ctypes.byref(h))
    assert rc == _lib.OQ_ERR_ARG and not h.value


def test_property_asserts(oq):
    with pytest.raises(AssertionError):
        oq.RateStateQuasiDynamicProperty(np.ones(3), np.ones(4), np.ones(3), np.ones(3), 1.0, 1.0)
    with pytest.raises(AssertionError):
        oq.RateStateQuasiDynamicProperty(np.ones(3), np.ones(3), np.ones(3), np.ones(3), 1.0, 1.0, 0.0)
    with pytest.raises(AssertionError):
        oq.PowerLawViscosityProperty(np.ones(3), np.ones(3), np.ones(5))
    with pytest.raises(AssertionError):
        oq.PowerLawViscosityProperty(np.ones(3), np.ones(2), np.ones(6))


def test_quadrature_format(oq):
    c, w = oq.get_quadrature("Gauss1")
    assert np.allclose(c, 0) and np.allclose(w, 1)
    c, w = oq.get_quadrature("Gauss2")
    assert len(w) == 8 and np.isclose(w.sum(), 1) and np.allclose(np.abs(c), 1 / np.sqrt(3))
    with pytest.raises(AssertionError):
        oq.get_quadrature((np.zeros(5), np.ones(2)))


def test_viscosity_law_host(oq, rng):
    # test/BEM/tests.jl:142-154
    A, n, s, t = rng.random(4)
    p = oq.PowerLawViscosityProperty([A], [n], rng.random(6))
    assert oq.dϵ_dt(p, 0, s, t) == A * s * t**n
    A2, n2 = rng.random(2)
    p2 = oq.PowerLawViscosityProperty([A2], [n2], rng.random(6))
    pc = oq.CompositePowerLawViscosityProperty([p, p2], rng.random(6))
    assert oq.dϵ_dt(pc, 0, s, t) == A * s * t**n + A2 * s * t**n2


def test_matvec_backend_preference(oq, caplog):
    # test/tests.jl:60-65
    assert oq.get_matvecmul() == "LinearAlgebra"
    assert oq.set_matvecmul("Octavian") == \
        "New backend Octavian set; restart your Julia session for this change to take effect!"
    assert oq.set_matvecmul("LinearAlgebra").startswith("New backend LinearAlgebra set")
    with pytest.raises(ValueError):
        oq.set_matvecmul("DummyMatVec")


def test_product_fails_loudly_without_library(oq, monkeypatch):
    from oetqf_amd import _lib
    monkeypatch.setattr(_lib, "LIBPATH", "/nonexistent/liboetqf_amd.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.OetqfError):
        _lib.load()


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_fft_grid_needs_a_grid_mantle(oq, rng):
    """OQ_K_FFT_GRID (K_mm as 2D convolutions over a mantle on a uniform
    horizontal grid) is refused before any device work: by oq_rhs_create,
    which has no mesh, and for a mantle refined along x ("Bump").  A box with
    graded layers passes the grid test (and then fails here only for want of
    a device)."""
    from helpers import rs_props, small_fault, visc_props
    from oetqf_amd import _lib
    lib = oq.load_library()
    a = np.ones(4)
    rs = _lib.RSProps(_lib.ptr(a), _lib.ptr(a), _lib.ptr(a), _lib.ptr(a), 1.0, 1e-9, 0.6, 1e-6)
    h = ctypes.c_void_p()
    st = np.zeros(8)
    rc = lib.oq_rhs_create(0, 2, 2, 0, ctypes.byref(rs), None, _lib.ptr(st), 0, None, None, None,
                           _lib.OQ_K_FFT_GRID, ctypes.byref(h))
    assert rc == _lib.OQ_ERR_ARG and not h.value
    assert "oq_rhs_create_from_mesh" in lib.oq_last_error().decode()
    mf = small_fault(90.0)
    pf = rs_props(mf, rng)

    def build(ma):
        pa = visc_props(ma, rng, 0.0)
        u = oq.ArrayPartition(np.ones((mf.nx, mf.nxi)), np.ones((mf.nx, mf.nxi)), np.zeros((ma.ne, 6)),
                              np.zeros((ma.ne, 6)), np.zeros((mf.nx, mf.nxi)))
        return oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u, (0.0, 1.0), kdtype="fft_grid")

    with pytest.raises(ValueError, match="uniform horizontal grid"):
        build(oq.hex8_box(-15e3, -5e3, -12e3, 30e3, 10e3, -20e3, 6, 2, 2, rfx=0.5))
    with pytest.raises(Exception) as ei:
        build(oq.hex8_box(-15e3, -5e3, -12e3, 30e3, 10e3, -20e3, 6, 2, 3, rfzh=[1.0, 1.5, 2.25]))
    assert "grid" not in str(ei.value)
    with pytest.raises(ValueError, match="kdtype"):
        oq.assemble(np.zeros((2, 2, 2)), pf, None, (0.0, 1.0), kdtype="fft_grid")
