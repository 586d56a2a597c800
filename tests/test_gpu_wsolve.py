"""wsolve (src/io.jl:118-133) on the device recorder, after the reference's
own storage test (test/tests.jl:1-42): the saved time series are the
accepted steps, every saved state is the integrator's state, append mode
concatenates, stride keeps every stride-th step, and an existing file is not
overwritten without force."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import rs_props, small_fault

pytestmark = pytest.mark.gpu
YR = 365 * 86400.0


def _prob(oracle, rng, tspan):
    mf = small_fault(90.0)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(0.1 * rng.standard_normal((mf.nx, mf.nxi)))
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros_like(v))
    return oq.assemble(st, pf, u0, tspan), mf


def getu(u, t, integrator):
    return (u.x[0], u.x[1], u.x[2])


def test_wsolve_matches_solve(oq, oracle, rng, tmp_path):
    prob, mf = _prob(oracle, rng, (0.0, 0.3 * YR))
    it = oq.Integrator(prob)
    it.record(step_capacity=100000)
    while it.step(4096) == "Running":
        pass
    step_t, _, _, _ = it.drain()
    tend, _, uend, na, nr = it.state()
    it.close()
    f = str(tmp_path / "out")
    sol = oq.wsolve(prob, oq.Tsit5(), f, 50, getu, ["u1", "u2", "u3"], "t")
    assert sol.retcode == "Success" and sol.naccept == na
    t = oq.wread(f, "t")
    assert np.array_equal(t, np.concatenate([[0.0], step_t]))
    u1 = oq.wread(f, "u1")
    assert u1.shape == (mf.nx, mf.nxi, na + 1)
    assert np.array_equal(u1[:, :, 0], prob.u0.x[0])
    assert np.array_equal(u1[:, :, -1], uend.x[0])
    assert np.array_equal(oq.wread(f, "u3")[:, :, -1], uend.x[2])

    # append: a second span continues the datasets (test/tests.jl:22-32)
    prob2, _ = _prob(oracle, np.random.default_rng(5), (0.4 * YR, 0.5 * YR))
    sol2 = oq.wsolve(prob2, oq.Tsit5(), f, 50, getu, ["u1", "u2", "u3"], "t", append=True)
    t2 = oq.wread(f, "t")
    assert len(t2) == len(t) + sol2.naccept + 1 and np.array_equal(t2[:len(t)], t)
    assert oq.wread(f, "u2").shape[-1] == len(t2)

    # stride (test/tests.jl:34-41): length(t) == naccept ÷ stride + 1
    sol3 = oq.wsolve(prob, oq.Tsit5(), f, 50, getu, ["u1", "u2", "u3"], "t", stride=11, force=True)
    t3 = oq.wread(f, "t")
    assert len(t3) == sol3.naccept // 11 + 1
    assert np.array_equal(t3, np.concatenate([[0.0], step_t])[::11])

    # no overwrite without force (io.jl:119-123)
    assert oq.wsolve(prob, oq.Tsit5(), f, 50, getu, ["u1", "u2", "u3"], "t") is None
