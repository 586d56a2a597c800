"""Edge shapes and special states against the oracle, for every operator
type of fault-only problems: a single patch, a single strike column or dip
row, ragged sizes that fill no tile exactly, and states holding the special
cases of update_fault! (theta <= 0, where max(0, theta) switches the state
term off, equation.jl:238; v = vpl, which contributes no stressing,
equation.jl:38)."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import relerr, rs_props

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 5), (7, 1), (33, 3), (65, 9)]
TOL = {"fp64": 1e-12, "toeplitz": 1e-12, "fft": 1e-12, "fp32": 1e-6}


@pytest.mark.parametrize("nx,nxi", SHAPES)
def test_edge_shapes_and_special_states(oracle, rng, nx, nxi):
    d = 1.25e3
    mf = oq.gen_mesh("RectOkada", nx * d, nxi * d, d, d, 60.0)
    assert (mf.nx, mf.nxi) == (nx, nxi)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(rng.standard_normal((nx, nxi)))
    th = pf.L / v * np.exp(0.3 * rng.standard_normal((nx, nxi)))
    v.flat[0] = pf.vpl  # no stressing from this patch
    if v.size > 1:
        th.flat[-1] = -abs(th.flat[-1])  # theta < 0
    if v.size > 2:
        th.flat[1] = -1e-3 * abs(th.flat[1])
    u = oq.ArrayPartition(v, th, rng.random((nx, nxi)))
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    ref = oracle.rhs_fault(pf, st, v, th)
    for kd, tol in TOL.items():
        prob = oq.assemble(st, pf, u, (0.0, 1.0), kdtype=kd)
        du = oq.similar(u)
        prob.f(du, u, None, 0.0)
        prob.f.close()
        for k in range(3):
            assert np.all(np.isfinite(du.x[k])), (kd, k)
            e = relerr(du.x[k], ref[k])
            assert e <= tol, (kd, k, e)
