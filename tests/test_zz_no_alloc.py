"""The allocation-free K_ff DFT builder (oq_gf_okada_ff_dft_noalloc, static
device scratch, device pointers in and out, caller's stream) against the
allocating entry point oq_gf_okada_ff_dft on the same meshes, through the C
ABI: two calls in a row on one device, the second smaller, each bit for bit
equal (the same kernels run on the same inputs)."""
import ctypes

import numpy as np
import pytest

import oetqf_amd as oq
from oetqf_amd import _lib
from oetqf_amd.greens import _Keep, fault_struct

LAM = MU = 3e10


def _reference(mf, ftype, nrept, beta):
    keep = _Keep()
    fs = fault_struct(mf, keep)
    out = np.zeros(2 * mf.nx * mf.nxi * mf.nxi)
    _lib.check(_lib.load().oq_gf_okada_ff_dft(ctypes.byref(fs), LAM, MU, ftype, nrept, beta,
                                              _lib.ptr(out)))
    return out


class _Hip:
    """The caller's side: device buffers and a stream from the HIP runtime
    the library itself is linked against (loaded with it)."""

    def __init__(self):
        _lib.load()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def _ok(self, err, what):
        assert err == 0, "%s failed: %d" % (what, err)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        p = ctypes.c_void_p()
        self._ok(self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))), "hipMalloc")
        self.bufs.append(p)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1),
                 "hipMemcpy H2D")
        return p

    def download(self, p, n):
        out = np.empty(n)
        self._ok(self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(8 * n), 2),
                 "hipMemcpy D2H")
        return out

    def stream(self):
        s = ctypes.c_void_p()
        self._ok(self.rt.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s

    def sync(self, s):
        self._ok(self.rt.hipStreamSynchronize(s), "hipStreamSynchronize")

    def close(self, s):
        self.rt.hipStreamDestroy(s)
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(hip, mf, ftype, nrept, beta, stream):
    arrs = [hip.upload(a) for a in (mf.x, mf.ax, mf.y, mf.z, mf.axi)]
    pd = ctypes.POINTER(ctypes.c_double)
    fs = _lib.FaultMesh(*[ctypes.cast(p, pd) for p in arrs], mf.nx, mf.nxi, mf.dx, mf.dep, mf.dip)
    n = 2 * mf.nx * mf.nxi * mf.nxi
    out = hip.upload(np.full(n, np.nan))
    print("noalloc call: %d x %d patches, ftype %d" % (mf.nx, mf.nxi, ftype), flush=True)
    rc = _lib.load().oq_gf_okada_ff_dft_noalloc(ctypes.byref(fs), LAM, MU, ftype, nrept, beta, out,
                                                stream)
    _lib.check(rc)
    hip.sync(stream)
    return hip.download(out, n)


def test_noalloc_rejects_oversize_before_any_device_call():
    """nx * nxi^2 above the 1048576-double scratch: OQ_ERR_ARG, checked
    before the device is touched (runs without a GPU)."""
    mf = oq.gen_mesh("RectOkada", 80e3, 20e3, 312.5, 250.0, 90.0)  # 256 x 80
    assert mf.nx * mf.nxi ** 2 > 1 << 20
    keep = _Keep()
    fs = fault_struct(mf, keep)
    rc = _lib.load().oq_gf_okada_ff_dft_noalloc(ctypes.byref(fs), LAM, MU, 0, 2, 0.0,
                                                ctypes.c_void_p(16), None)
    assert rc == _lib.OQ_ERR_ARG
    assert "scratch" in _lib.load().oq_last_error().decode()


@pytest.mark.gpu
def test_noalloc_dft_equals_allocating_builder():
    hip = _Hip()
    stream = hip.stream()
    big = oq.gen_mesh_graded(30e3, 10e3, 1.25e3, 10, 60.0, ratio=1.2)   # 24 x 10
    small = oq.gen_mesh("RectOkada", 15e3, 5e3, 1.25e3, 1.25e3, 90.0)   # 12 x 4
    try:
        for mf, ftype, nrept, beta in ((big, 1, 2, 0.5), (small, 0, 1, 0.0)):
            got = _noalloc(hip, mf, ftype, nrept, beta, stream)
            want = _reference(mf, ftype, nrept, beta)
            assert np.all(np.isfinite(got))
            assert np.array_equal(got, want), np.abs(got - want).max()
    finally:
        hip.close(stream)
