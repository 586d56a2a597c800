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


def _coupled_prob(oracle, rng, tspan):
    from helpers import small_mantle, visc_props
    mf, ma = small_fault(90.0), small_mantle()
    lam = mu = 3e10
    pf = rs_props(mf, rng)
    pa = visc_props(ma, rng, 0.0)
    v = pf.vpl * np.exp(0.1 * rng.standard_normal((mf.nx, mf.nxi)))
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros((ma.ne, 6)), 1e6 * rng.standard_normal((ma.ne, 6)),
                           np.zeros_like(v))
    gfs = (oracle.gf_ff(mf, lam, mu, buffer_ratio=1.0), oracle.gf_mf(mf, ma, lam, mu, buffer_ratio=1.0),
           oracle.gf_fm(ma, mf, lam, mu), oracle.gf_mm(ma, lam, mu))
    return oq.assemble(*gfs, pf, pa, u0, tspan), gfs, pf, pa


def test_wsolve_vcabm5_matches_solve(oq, oracle, rng, tmp_path):
    """wsolve(prob, VCABM5(), ...) integrates with VCABM5: its final state and
    step counts equal solve(prob, VCABM5())."""
    prob, mf = _prob(oracle, rng, (0.0, 0.2 * YR))
    sol = oq.solve(prob, oq.VCABM5())
    f = str(tmp_path / "ab")
    w = oq.wsolve(prob, oq.VCABM5(), f, 64, oq.VΘΔ, ["v", "θ", "δ"], "t")
    assert w.retcode == sol.retcode == "Success"
    assert (w.naccept, w.nreject) == (sol.naccept, sol.nreject)
    assert np.array_equal(w.u.pack(), sol.u.pack())
    assert np.array_equal(oq.wread(f, "v")[:, :, -1], sol.u.x[0])
    assert len(oq.wread(f, "t")) == sol.naccept + 1
    with pytest.raises(NotImplementedError):
        oq.wsolve(prob, object(), f, 64, oq.VΘΔ, ["v", "θ", "δ"], "t", force=True)


@pytest.mark.parametrize("alg", ["tsit5", "vcabm5"])
def test_wsolve_derivative_handler(oq, oracle, rng, tmp_path, alg):
    """The reference example's handler (examples/otf-with-mantle.jl:157-159;
    src/io.jl:86) saves integrator(integrator.t, Val{1}).x[3], the strain
    rate f(u) at each saved state: every saved strain rate must equal the
    oracle's RHS evaluated on the saved stress."""
    prob, gfs, pf, pa = _coupled_prob(oracle, rng, (0.0, 0.05 * YR))
    handler = lambda u, t, it: (u.x[0], u.x[1], it(it.t, oq.Val[1]).x[2], u.x[2], u.x[3], u.x[4])
    f = str(tmp_path / "ex")
    a = oq.VCABM5() if alg == "vcabm5" else oq.Tsit5()
    sol = oq.wsolve(prob, a, f, 100, handler, ["v", "θ", "dϵ", "ϵ", "σ", "δ"], "t", stride=7)
    assert sol.retcode == "Success"
    de, sg, v, th = (oq.wread(f, k) for k in ("dϵ", "σ", "v", "θ"))
    nt = de.shape[-1]
    assert nt == sol.naccept // 7 + 1 and sg.shape[-1] == nt
    worst = 0.0
    for k in range(nt):
        ref = oracle.rhs_coupled(pf, pa, *gfs, v[..., k], th[..., k], sg[..., k])
        worst = max(worst, np.abs(de[..., k] - ref[2]).max() / np.abs(ref[2]).max())
    assert worst <= 1e-12, worst
    # the two-part handler of io.jl:86 on the same run
    g = str(tmp_path / "ex2")
    oq.wsolve(prob, a, g, 100, oq.VΘdεΔ, ["v", "θ", "dϵ", "δ"], "t", stride=7)
    assert np.array_equal(oq.wread(g, "dϵ"), de)


@pytest.mark.parametrize("alg", ["tsit5", "vcabm5"])
def test_wsolve_persistent_derivative(oq, oracle, rng, tmp_path, alg):
    """The persistent steppers record f(u) with each snapshot too: the saved
    dv/dt equals the oracle's fault-only RHS on the saved state."""
    mf = small_fault(90.0)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(0.1 * rng.standard_normal((mf.nx, mf.nxi)))
    prob = oq.assemble(st, pf, oq.ArrayPartition(v, pf.L / v, np.zeros_like(v)), (0.0, 0.1 * YR))
    f = str(tmp_path / "p")
    a = oq.VCABM5() if alg == "vcabm5" else oq.Tsit5()
    h = lambda u, t, it: (u.x[0], u.x[1], it(it.t, oq.Val[1]).x[0])
    sol = oq.wsolve(prob, a, f, 64, h, ["v", "θ", "dv"], "t", stride=5, persistent=True)
    assert sol.retcode == "Success"
    v, th, dv = (oq.wread(f, k) for k in ("v", "θ", "dv"))
    assert dv.shape[-1] == sol.naccept // 5 + 1
    worst = 0.0
    for k in range(dv.shape[-1]):
        ref = oracle.rhs_fault(pf, st, v[..., k], th[..., k])[0]
        worst = max(worst, np.abs(dv[..., k] - ref).max() / np.abs(ref).max())
    assert worst <= 1e-9, worst


def test_recorder_overflow_then_next_drain(oq, oracle, rng):
    """More accepted steps than the rings hold: the drain hands out the
    newest records in order with a warning naming the loss, and the next
    drain succeeds."""
    mf = small_fault(90.0)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(3.0 * rng.standard_normal((mf.nx, mf.nxi)))  # far from steady state
    prob = oq.assemble(oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0), pf,
                       oq.ArrayPartition(v, pf.L / v, np.zeros_like(v)), (0.0, 30 * YR))
    it = oq.Integrator(prob)
    it.record(stride=1, snap_capacity=8, step_capacity=16)
    it.step(400)
    _, _, _, na0, _ = it.state()
    assert na0 > 16, na0
    with pytest.warns(RuntimeWarning, match="overflow"):
        st_t, st_v, sn_t, sn_u = it.drain()
    assert len(st_t) == 16 and len(sn_t) == 8
    assert np.all(np.diff(st_t) > 0) and np.array_equal(sn_t, st_t[-8:])
    t, _, u, _, _ = it.state()
    assert st_t[-1] == t and np.array_equal(sn_u[-1], u.pack())
    it.step(5)
    st_t2, _, sn_t2, _ = it.drain()  # no warning
    assert 0 < len(st_t2) <= 5 and st_t2[0] > st_t[-1]
    it.close()
