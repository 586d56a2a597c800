"""GPU Green's-function builders vs the CPU oracle (FP64, <= 1e-10 relative).

Reference: src/BEM/GF.jl:31-71 (K_ff), :123-174 (K_mf), :194-227 (K_fm),
:250-296 (K_mm); the GPU path is called through the C ABI.
"""
import numpy as np
import pytest

from helpers import maxrel, relerr, small_fault, small_mantle

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.mark.parametrize("dip", [90.0, 41.0])
@pytest.mark.parametrize("ftype", ["ss", "ds"])
@pytest.mark.parametrize("buffer_ratio", [0.0, 1.0])
def test_gf_ff_matches_oracle(oq, oracle, dip, ftype, buffer_ratio):
    mf = small_fault(dip)
    ft = oq.StrikeSlip() if ftype == "ss" else oq.DipSlip()
    g = oq.stress_greens_function(mf, 3e10, 3e10, ftype=ft, fourier=False, buffer_ratio=buffer_ratio)
    r = oracle.gf_ff(mf, 3e10, 3e10, ftype=ft.code, nrept=2, buffer_ratio=buffer_ratio)
    assert g.shape == (mf.nx, mf.nxi, mf.nxi)
    assert relerr(g, r) <= TOL and maxrel(g, r) <= TOL


def test_gf_ff_fourier_is_rfft_of_circulant(oq, oracle):
    """fourier=true returns rfft([st; reverse(st[2:end])]) (GF.jl:60-67)."""
    mf = small_fault(41.0)
    g = oq.stress_greens_function(mf, 3e10, 3e10, fourier=True)
    st = oracle.gf_ff(mf, 3e10, 3e10)
    circ = np.concatenate([st, st[:0:-1]], axis=0)
    ref = np.fft.rfft(circ, axis=0)
    assert g.shape == ref.shape == (mf.nx, mf.nxi, mf.nxi)
    assert relerr(g, ref) <= TOL


def test_gf_ff_fourier_at_scale(oq, oracle):
    """The device DFT (one workgroup per (j, l) column) on the 256 x 64 fault
    of the N = 65536 workload: numpy's rfft of the circulant of the oracle's
    kernel, to the Green's tolerance; its imaginary parts are exactly 0."""
    mf = oq.gen_mesh("RectOkada", 80e3, 20e3, 312.5, 312.5, 90.0)
    g = oq.stress_greens_function(mf, 3e10, 3e10, fourier=True)
    st = oracle.gf_ff(mf, 3e10, 3e10)
    ref = np.fft.rfft(np.concatenate([st, st[:0:-1]], axis=0), axis=0)
    assert relerr(g, ref) <= TOL and np.all(g.imag == 0.0)
    assert np.abs(ref.imag).max() <= 1e-12 * np.abs(ref.real).max()


@pytest.mark.parametrize("qtype", ["Gauss1", "Gauss2"])
@pytest.mark.parametrize("ftype", ["ss", "ds"])
def test_gf_mf_matches_oracle(oq, oracle, qtype, ftype):
    mf, ma = small_fault(90.0), small_mantle()
    ft = oq.StrikeSlip() if ftype == "ss" else oq.DipSlip()
    g = oq.stress_greens_function(mf, ma, 3e10, 3e10, ftype=ft, qtype=qtype, buffer_ratio=1.0)
    qc, qw = oq.get_quadrature(qtype)
    r = oracle.gf_mf(mf, ma, 3e10, 3e10, ftype=ft.code, qc=qc, qw=qw, nrept=2, buffer_ratio=1.0)
    assert g.shape == (6 * ma.ne, mf.nx * mf.nxi)
    assert relerr(g, r) <= TOL and maxrel(g, r) <= TOL


@pytest.mark.parametrize("dip", [90.0, 60.0])
@pytest.mark.parametrize("ftype", ["ss", "ds"])
def test_gf_fm_matches_oracle(oq, oracle, dip, ftype):
    mf, ma = small_fault(dip), small_mantle()
    ft = oq.StrikeSlip() if ftype == "ss" else oq.DipSlip()
    g = oq.stress_greens_function(ma, mf, 3e10, 3e10, ftype=ft)
    r = oracle.gf_fm(ma, mf, 3e10, 3e10, ftype=ft.code, nquad=6)
    assert g.shape == (mf.nx * mf.nxi, 6 * ma.ne)
    assert relerr(g, r) <= TOL and maxrel(g, r) <= TOL


@pytest.mark.parametrize("qtype", ["Gauss1", "Gauss2"])
def test_gf_mm_matches_oracle(oq, oracle, qtype):
    ma = small_mantle()
    g = oq.stress_greens_function(ma, 3e10, 3e10, qtype=qtype, checkeigvals=False)
    qc, qw = oq.get_quadrature(qtype)
    r = oracle.gf_mm(ma, 3e10, 3e10, qc=qc, qw=qw, nquad=6)
    assert g.shape == (6 * ma.ne, 6 * ma.ne)
    assert relerr(g, r) <= TOL and maxrel(g, r) <= TOL


def test_gf_receivers_on_cell_faces(oq, oracle):
    """Fault receivers on the y = 0 faces of mantle cells (the composite
    geometry): the traction components are finite and match the oracle."""
    mf = oq.gen_mesh("RectOkada", 20e3, 16e3, 2.5e3, 2e3, 90.0)
    ma = oq.hex8_box(-10e3, -5e3, -8e3, 20e3, 10e3, -12e3, 4, 4, 2)  # y faces at -5, -2.5, 0, 2.5, 5 km
    g = oq.stress_greens_function(ma, mf, 3e10, 3e10)
    r = oracle.gf_fm(ma, mf, 3e10, 3e10, nquad=6)
    assert np.all(np.isfinite(g))
    assert relerr(g, r) <= TOL
