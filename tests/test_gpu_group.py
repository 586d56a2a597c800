"""Grouped batches of uneven geometries (FP32 operators, 16-member MFMA
tiles, one grouped launch per tile shape): the RHS against the FP64 oracle
member by member, exactness against the rounded operator, invariance under
the autotuned tile shapes, and the lockstep grouped Tsit5 against
single-member runs."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import relerr, small_fault, small_mantle, visc_props

pytestmark = pytest.mark.gpu
LAM = MU = 3e10


def _geoms(rng):
    """Three geometries of uneven size: 32 patches (fewer rows than one
    tile), 100 patches + 12 cells, 900 patches."""
    g0 = (small_fault(90.0), None, None)
    mf1, ma1 = small_fault(90.0, x=40e3, xi=10e3, d=2e3), small_mantle()
    g1 = (mf1, ma1, visc_props(ma1, rng, 0.0))
    g2 = (small_fault(90.0, x=60e3, xi=15e3, d=1e3), None, None)
    return [g0, g1, g2]


def _members(mf, rng, M):
    shp = (mf.nx, mf.nxi)
    out = []
    for _ in range(M):
        a = 0.015 + 0.002 * rng.random(shp)
        b = a + rng.uniform(0.002, 0.006)
        out.append(oq.RateStateQuasiDynamicProperty(a, b, np.full(shp, rng.uniform(4e-3, 4e-2)),
                                                    np.full(shp, rng.uniform(30e6, 80e6)),
                                                    MU / (2 * 3040.0), 1e-9 * (1 + 0.1 * rng.random())))
    return out


def _state(mf, ma, p, rng):
    shp = (mf.nx, mf.nxi)
    v = p.vpl * np.exp(0.5 * rng.standard_normal(shp))
    th = p.L / v * np.exp(0.2 * rng.standard_normal(shp))
    if ma is None:
        return oq.ArrayPartition(v, th, rng.random(shp))
    return oq.ArrayPartition(v, th, 1e-6 * rng.standard_normal((ma.ne, 6)),
                             1e6 * rng.standard_normal((ma.ne, 6)), rng.random(shp))


@pytest.fixture(scope="module")
def batch():
    rng = np.random.default_rng(41)
    geoms = _geoms(rng)
    M = 5
    members = [_members(mf, rng, M) for mf, _, _ in geoms]
    states = [[_state(mf, ma, p, rng) for p in mem] for (mf, ma, _), mem in zip(geoms, members)]
    grp = oq.Group(geoms, LAM, MU, members, buffer_ratio=1.0)
    U = grp.pack([np.stack([s.pack() for s in st]) for st in states])
    yield geoms, members, states, grp, U
    grp.close()


def test_group_rhs_matches_fp64_oracle(oracle, batch):
    geoms, members, states, grp, U = batch
    dU = grp.unpack(grp.eval(U))
    worst = 0.0
    for g, (mf, ma, pa) in enumerate(geoms):
        st = oracle.gf_ff(mf, LAM, MU, buffer_ratio=1.0)
        if ma is not None:
            g12 = oracle.gf_mf(mf, ma, LAM, MU, buffer_ratio=1.0)
            g21 = oracle.gf_fm(ma, mf, LAM, MU)
            g22 = oracle.gf_mm(ma, LAM, MU)
        for m, (p, s) in enumerate(zip(members[g], states[g])):
            if ma is None:
                ref = oracle.rhs_fault(p, st, s.x[0], s.x[1])
            else:
                ref = oracle.rhs_coupled(p, pa, st, g12, g21, g22, s.x[0], s.x[1], s.x[3])
            got = s.similar().unpack(dU[g][m])
            for k in range(len(ref)):
                e = relerr(got.x[k], ref[k])
                worst = max(worst, e)
                assert e <= 1e-5, "geometry %d member %d component %d: %g" % (g, m, k, e)
    print("grouped RHS worst rel err vs FP64 oracle: %.2e" % worst)


def test_group_rhs_exact_on_rounded_operator(oracle, batch):
    """Against the FP64 oracle applied to the FP32 operator itself the error
    is the FP32 product/partial-sum rounding only."""
    geoms, members, states, grp, U = batch
    dU = grp.unpack(grp.eval(U))
    mf, ma, pa = geoms[1]
    nf = mf.nx * mf.nxi
    for m in (0, 4):
        p, s = members[1][m], states[1][m]
        prob = oq.assemble_from_mesh(mf, ma, LAM, MU, p, pa, s, (0.0, 1.0), buffer_ratio=1.0,
                                     kdtype="fp32")
        K = prob.f.dense_operator()
        prob.f.close()
        ref = oracle.rhs_dense(p, pa, K, nf, ma.ne, s.x[0].reshape(-1, order="F"),
                               s.x[1].reshape(-1, order="F"), s.x[3].reshape(-1, order="F"))
        got = s.similar().unpack(dU[1][m])
        assert relerr(got.x[0].reshape(-1, order="F"), ref[0]) <= 2e-6
        assert relerr(got.x[3].reshape(-1, order="F"), ref[3]) <= 2e-6


def test_group_autotune_keeps_results(batch):
    geoms, members, states, grp, U = batch
    before = grp.eval(U)
    choice = grp.autotune(iters=3)
    print("autotuned tile shapes:", choice)
    assert set(choice) == {0}  # all three geometries have N < 2048
    after = grp.eval(U)
    assert relerr(after, before) <= 1e-6


def test_group_every_tile_shape_gives_the_same_rhs(batch):
    """Each of the tile shapes (register-streaming, LDS-staged with one or
    two slices in flight), pinned for every bucket, reproduces the default
    shape's RHS."""
    from oetqf_amd.group import SHAPES
    geoms, members, states, grp, U = batch
    ref = grp.eval(U)
    for s in SHAPES:
        grp.set_shapes({b: s for b in range(8)})
        e = relerr(grp.eval(U), ref)
        print("%-11s rel diff %.2e" % (s, e))
        assert e <= 1e-6, s


@pytest.mark.parametrize("alg", ["tsit5", "vcabm5"])
def test_group_stepper_matches_single_members(oq, alg):
    rng = np.random.default_rng(43)
    geoms = _geoms(rng)
    M = 2
    members = [_members(mf, rng, M) for mf, _, _ in geoms]
    U0 = []
    for (mf, ma, _), mem in zip(geoms, members):
        rows = []
        for p in mem:
            v = p.vpl * np.exp(0.1 * rng.standard_normal((mf.nx, mf.nxi)))
            parts = [v, p.L / v]
            if ma is not None:
                parts += [np.zeros((ma.ne, 6)), 1e5 * np.ones((ma.ne, 6))]
            parts.append(np.zeros((mf.nx, mf.nxi)))
            rows.append(oq.ArrayPartition(*parts))
        U0.append(rows)
    T = 0.2 * 365 * 86400.0
    grp = oq.Group(geoms, LAM, MU, members, buffer_ratio=1.0)
    method = oq.VCABM5() if alg == "vcabm5" else oq.Tsit5()
    grp.start(grp.pack([np.stack([u.pack() for u in r]) for r in U0]), (0.0, T), dt0=1.0, alg=method)
    while grp.run(256) > 0:
        pass
    t, U, na, nr, rc = grp.state()
    grp.close()
    assert all(r == "Success" for r in rc) and np.all(t == T)
    Us = grp.unpack(U)
    for g, (mf, ma, pa) in enumerate(geoms):
        nf = mf.nx * mf.nxi
        for m, p in enumerate(members[g]):
            prob = oq.assemble_from_mesh(mf, ma, LAM, MU, p, pa, U0[g][m], (0.0, T),
                                         buffer_ratio=1.0, kdtype="fp32")
            sol = oq.solve(prob, method, reltol=1e-6, abstol=1e-8, dt=1.0)
            ref = sol.u.pack()
            assert relerr(Us[g][m][:nf], ref[:nf]) <= 1e-4
            assert relerr(Us[g][m][nf:2 * nf], ref[nf:2 * nf]) <= 1e-5


def test_group_checkpoint_resume_is_the_same_run(oq):
    """Checkpoint after some lockstep steps, resume in a second group with the
    same tile shapes: the continuation is bit for bit the uninterrupted run."""
    rng = np.random.default_rng(47)
    geoms = _geoms(rng)
    members = [_members(mf, rng, 2) for mf, _, _ in geoms]
    states = [np.stack([_state(mf, ma, p, rng).pack() for p in mem]) for (mf, ma, _), mem in zip(geoms, members)]
    T = 0.3 * 365 * 86400.0
    a = oq.Group(geoms, LAM, MU, members, buffer_ratio=1.0)
    U0 = a.pack(states)
    a.start(U0, (0.0, T), dt0=1.0)
    a.run(8)  # a few lockstep attempts: every member is still running
    U, ctrl = a.checkpoint()
    assert np.all(ctrl["naccept"] > 0) and np.all(ctrl["t"] < T)
    while a.run(256) > 0:
        pass
    t1, U1, na1, nr1, rc1 = a.state()
    a.close()
    b = oq.Group(geoms, LAM, MU, members, buffer_ratio=1.0)
    b.resume(U, ctrl, T)
    while b.run(256) > 0:
        pass
    t2, U2, na2, nr2, rc2 = b.state()
    b.close()
    assert rc1 == rc2 and all(r == "Success" for r in rc1)
    assert np.array_equal(t1, t2) and np.array_equal(U1, U2)
    assert np.array_equal(na1, na2) and np.array_equal(nr1, nr2)
