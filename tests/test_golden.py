"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from
the pinned oracle): the oracle must keep reproducing them, and the GPU
builders must match them."""
import json
import os

import numpy as np
import pytest

from helpers import relerr

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def test_oracle_reproduces_dc3d_golden(oracle):
    for c in load("dc3d.json"):
        u, ret = oracle.dc3d(c["x"], c["y"], c["z"], c["alpha"], c["depth"], c["dip"], *c["al"],
                             *c["aw"], *c["slip"])
        assert ret == c["ret"]
        assert np.allclose(u, c["u"], rtol=1e-13, atol=1e-13 * np.abs(c["u"]).max())


def test_oracle_reproduces_hex8_golden(oracle):
    for c in load("hex8.json"):
        s = oracle.stress_vol_hex8(c["x"], c["y"], c["z"], *c["q"], c["T"], c["L"], c["W"],
                                   np.array(c["eps"]), c["mu"], c["nu"], 6)
        assert np.allclose(s, c["sigma"], rtol=1e-12, atol=1e-12 * max(1.0, np.abs(c["sigma"]).max()))


def test_oracle_reproduces_kff_golden(oracle, oq):
    c = load("kff.json")
    mf = oq.gen_mesh("RectOkada", *c["mesh"])
    st = oracle.gf_ff(mf, c["lam"], c["mu"], ftype=c["ftype"], nrept=c["nrept"],
                      buffer_ratio=c["buffer_ratio"])
    assert relerr(st.reshape(-1, order="F"), np.array(c["st_colmajor"])) <= 1e-14


@pytest.mark.gpu
def test_gpu_kff_matches_golden(oq):
    c = load("kff.json")
    mf = oq.gen_mesh("RectOkada", *c["mesh"])
    st = oq.stress_greens_function(mf, c["lam"], c["mu"], fourier=False, nrept=c["nrept"],
                                   buffer_ratio=c["buffer_ratio"])
    assert relerr(st.reshape(-1, order="F"), np.array(c["st_colmajor"])) <= 1e-10
