"""Ensemble RHS (one MFMA GEMM over a shared FP32/BF16 operator, or the
split-FP16 operator on the FP16 matrix cores) and the lockstep ensemble
Tsit5, against the FP64 oracle member by member."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import relerr, small_fault, small_mantle, visc_props

pytestmark = pytest.mark.gpu


def _members(mf, rng, M):
    shp = (mf.nx, mf.nxi)
    out = []
    for _ in range(M):
        a = 0.015 + 0.002 * rng.random(shp)
        b = a + rng.uniform(0.002, 0.006)
        out.append(oq.RateStateQuasiDynamicProperty(a, b, np.full(shp, rng.uniform(4e-3, 4e-2)),
                                                    np.full(shp, rng.uniform(30e6, 80e6)),
                                                    30e9 / (2 * 3040.0), 1e-9 * (1 + 0.1 * rng.random())))
    return out


def _state(mf, ma, p, rng):
    v = p.vpl * np.exp(0.5 * rng.standard_normal((mf.nx, mf.nxi)))
    th = p.L / v * np.exp(0.2 * rng.standard_normal((mf.nx, mf.nxi)))
    return oq.ArrayPartition(v, th, 1e-6 * rng.standard_normal((ma.ne, 6)),
                             1e6 * rng.standard_normal((ma.ne, 6)), rng.random((mf.nx, mf.nxi)))


@pytest.mark.parametrize("kdtype,tol,M", [("fp32", 1e-5, 70), ("bf16", 2e-2, 70), ("fp16x2", 1e-5, 70),
                                          ("fp16x2", 1e-5, 260)])
def test_ensemble_rhs_matches_fp64_oracle(oq, oracle, rng, kdtype, tol, M):
    """M = 70 spans two member tiles of 64 (FP32/BF16); M = 260 two of 256
    (split FP16)."""
    mf, ma = small_fault(90.0), small_mantle()
    lam = mu = 3e10
    members = _members(mf, rng, M)
    pa = visc_props(ma, rng, 0.0)
    states = [_state(mf, ma, p, rng) for p in members]
    ens = oq.Ensemble(mf, ma, lam, mu, members, pa, buffer_ratio=1.0, kdtype=kdtype)
    dU = ens.eval(np.stack([s.pack() for s in states]))
    st = oracle.gf_ff(mf, lam, mu, buffer_ratio=1.0)
    g12 = oracle.gf_mf(mf, ma, lam, mu, buffer_ratio=1.0)
    g21 = oracle.gf_fm(ma, mf, lam, mu)
    g22 = oracle.gf_mm(ma, lam, mu)
    worst = 0.0
    for m in sorted({0, 1, 33, 64, 69, M - 5, M - 1}):
        s = states[m]
        ref = oracle.rhs_coupled(members[m], pa, st, g12, g21, g22, s.x[0], s.x[1], s.x[3])
        got = s.similar().unpack(dU[m])
        for k in range(5):
            e = relerr(got.x[k], ref[k])
            worst = max(worst, e)
            assert e <= tol, "member %d component %d: %g" % (m, k, e)
    print("%s ensemble RHS worst rel err vs FP64 oracle: %.2e" % (kdtype, worst))


@pytest.mark.parametrize("alg", ["tsit5", "vcabm5"])
@pytest.mark.parametrize("kdtype", ["fp32", "fp16x2"])
def test_ensemble_stepper_matches_single_member(oq, rng, kdtype, alg):
    mf, ma = small_fault(90.0), small_mantle()
    members = _members(mf, rng, 3)
    pa = visc_props(ma, rng, 0.0)
    U0 = []
    for p in members:
        v = p.vpl * np.exp(0.1 * rng.standard_normal((mf.nx, mf.nxi)))
        U0.append(oq.ArrayPartition(v, p.L / v, np.zeros((ma.ne, 6)), 1e5 * np.ones((ma.ne, 6)),
                                    np.zeros((mf.nx, mf.nxi))))
    T = 0.2 * 365 * 86400.0
    ens = oq.Ensemble(mf, ma, 3e10, 3e10, members, pa, buffer_ratio=1.0, kdtype=kdtype)
    method = oq.VCABM5() if alg == "vcabm5" else oq.Tsit5()
    ens.start(np.stack([u.pack() for u in U0]), (0.0, T), reltol=1e-6, abstol=1e-8, dt0=1.0, alg=method)
    while ens.run(256) > 0:
        pass
    t, U, na, nr, rc = ens.state()
    assert all(r == "Success" for r in rc) and np.all(t == T)
    for m, p in enumerate(members):
        prob = oq.assemble_from_mesh(mf, ma, 3e10, 3e10, p, pa, U0[m], (0.0, T), buffer_ratio=1.0,
                                     kdtype="fp32")
        sol = oq.solve(prob, method, reltol=1e-6, abstol=1e-8, dt=1.0)
        if alg == "vcabm5":  # the same method and step sequence up to FP32 rounding of the products
            assert abs(int(na[m]) - sol.naccept) <= max(2, sol.naccept // 50), (na[m], sol.naccept)
        ref = sol.u.pack()
        nf = mf.nx * mf.nxi
        assert relerr(U[m][:nf], ref[:nf]) <= 1e-4       # slip rate
        assert relerr(U[m][nf:2 * nf], ref[nf:2 * nf]) <= 1e-5  # state variable


@pytest.mark.parametrize("kdtype", ["fp32", "fp16x2"])
def test_ensemble_checkpoint_resume_is_the_same_run(oq, rng, kdtype):
    """A lockstep run stopped after some steps, checkpointed, and resumed in a
    second ensemble on the same operator continues bit for bit as the
    uninterrupted run: states, times and every step counter."""
    mf, ma = small_fault(90.0), small_mantle()
    members = _members(mf, rng, 3)
    pa = visc_props(ma, rng, 0.0)
    U0 = np.stack([_state(mf, ma, p, rng).pack() for p in members])
    T = 0.3 * 365 * 86400.0
    kw = dict(buffer_ratio=1.0, kdtype=kdtype)
    a = oq.Ensemble(mf, ma, 3e10, 3e10, members, pa, **kw)
    a.start(U0, (0.0, T), dt0=1.0)
    a.run(8)  # a few lockstep attempts: every member is still running
    U, ctrl = a.checkpoint()
    assert np.all(ctrl["naccept"] > 0) and np.all(ctrl["t"] < T)
    while a.run(256) > 0:
        pass
    t1, U1, na1, nr1, rc1 = a.state()
    c1 = a.checkpoint()[1]
    a.close()
    b = oq.Ensemble(mf, ma, 3e10, 3e10, members, pa, **kw)
    b.resume(U, ctrl, T)
    while b.run(256) > 0:
        pass
    t2, U2, na2, nr2, rc2 = b.state()
    c2 = b.checkpoint()[1]
    b.close()
    assert rc1 == rc2 and all(r == "Success" for r in rc1)
    assert np.array_equal(t1, t2) and np.array_equal(U1, U2)
    assert np.array_equal(na1, na2) and np.array_equal(nr1, nr2)
    for k in ("dt", "qold", "iter"):
        assert np.array_equal(c1[k], c2[k]), k
