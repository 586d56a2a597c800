"""The composite, ensemble and grouped workloads at their real sizes, against
the oracle.

- Composite model: 128x32 transform fault (rows graded x1.05) + 16x16x8 hex8
  mantle, N = 16384, K assembled on the device in FP64.  The oracle rebuilds
  the rows of sampled patches above the mantle (z > -10 km) and on the
  y = 0 cell faces inside it (z < -10 km), and all six stress rows of
  sampled cells next to the fault and away from it; the device RHS rows
  must agree to 1e-10 (two independent FP64 assemblies).
- 256-member ensemble on that geometry, FP32 and split-FP16 operators:
  sampled members' rows against the oracle to 1e-5.
- Grouped batch: 24 geometries (1024-8192 patches, half with mantles) x 16
  members, FP32 operators: sampled member rows of the smallest and the
  largest geometry against the oracle to 1e-5.
"""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import relerr
from oetqf_amd.workloads import ensemble_members, grouped_batch, transform_fault_composite
from oracle_rows import OracleRows

pytestmark = pytest.mark.gpu


def _perturbed(pf, ma, rng):
    """An interseismic-like state: slip rates over four decades around vpl,
    theta around steady state, mantle stresses of 1 MPa."""
    shp = pf.a.shape
    v = pf.vpl * np.exp(2.0 * rng.standard_normal(shp))
    th = pf.L / pf.vpl * np.exp(rng.standard_normal(shp))
    parts = [v, th]
    if ma is not None:
        parts += [np.zeros((ma.ne, 6)), 1e6 * rng.standard_normal((ma.ne, 6))]
    parts.append(rng.random(shp))
    return oq.ArrayPartition(*parts)


def _composite_samples(w):
    mf, ma = w["mf"], w["ma"]
    J_off = np.flatnonzero(mf.z > -10e3)
    J_on = np.flatnonzero(mf.z < -10e3)
    J = np.unique(np.concatenate([J_off[[0, len(J_off) // 2, -1]], J_on[[0, len(J_on) // 2, -1]]]))
    I = np.array([0, 1, mf.nx // 2 - 1, mf.nx // 2, mf.nx - 1])
    near = np.flatnonzero(np.abs(ma.cy) < 0.6 * ma.dy)
    C = np.unique(np.concatenate([[0, ma.ne - 1], near[[0, len(near) // 3, -1]]]))
    return I, J, C


@pytest.fixture(scope="module")
def composite(oracle):
    w = transform_fault_composite()
    I, J, C = _composite_samples(w)
    print("composite: N = %d; oracle rows I=%s J=%s C=%s" % (w["mf"].nx * w["mf"].nxi + 6 * w["ma"].ne,
                                                           I, J, C), flush=True)
    rows = OracleRows(oracle, w["mf"], w["ma"], w["lam"], w["mu"], I, J, C)
    return w, rows


def test_composite_fp64_rows_match_oracle(composite):
    w, rows = composite
    u = _perturbed(w["pf"], w["ma"], np.random.default_rng(11))
    prob = oq.assemble_from_mesh(w["mf"], w["ma"], w["lam"], w["mu"], w["pf"], w["pa"], w["u0"],
                                 w["tspan"])
    try:
        du = oq.similar(u)
        prob.f(du, u, None, 0.0)
    finally:
        prob.f.close()
    dv_ref, ds_ref = rows.rows(w["pf"], w["pa"], u.x[0], u.x[1], u.x[3])
    dv, ds = rows.pick(du.x[0], du.x[3])
    e_dv, e_ds = relerr(dv, dv_ref), relerr(ds, ds_ref)
    # rows on the cell faces alone
    on = np.flatnonzero(np.repeat(w["mf"].z[rows.J] < -10e3, len(rows.I)))
    e_on = relerr(dv[on], dv_ref[on])
    print("composite FP64 rows vs oracle: dv %.2e (on cell faces %.2e)  dsigma %.2e"
          % (e_dv, e_on, e_ds), flush=True)
    assert e_dv <= 1e-10 and e_on <= 1e-10 and e_ds <= 1e-10


def test_composite_fft_grid_matches_dense_and_oracle(composite):
    """The composite model with K_ff by DFT and K_mm as 2D convolutions over
    the 16x16x8 mantle grid: the sampled rows against the oracle to 1e-10,
    every component against the dense FP64 model to 1e-12, in 1/2.4 of the
    operator bytes."""
    w, rows = composite
    u = _perturbed(w["pf"], w["ma"], np.random.default_rng(12))
    args = (w["mf"], w["ma"], w["lam"], w["mu"], w["pf"], w["pa"], w["u0"], w["tspan"])
    pg = oq.assemble_from_mesh(*args, kdtype="fft_grid")
    pd = oq.assemble_from_mesh(*args)
    try:
        dg, dd = oq.similar(u), oq.similar(u)
        pg.f(dg, u, None, 0.0)
        pd.f(dd, u, None, 0.0)
        kg, kd = pg.f.info()[2], pd.f.info()[2]
    finally:
        pg.f.close()
        pd.f.close()
    dv_ref, ds_ref = rows.rows(w["pf"], w["pa"], u.x[0], u.x[1], u.x[3])
    dv, ds = rows.pick(dg.x[0], dg.x[3])
    errs = [relerr(dg.x[k], dd.x[k]) for k in range(5)]
    print("composite fft_grid: operator %.3f GB (dense %.3f GB); vs dense %s; rows vs oracle dv %.2e "
          "dsigma %.2e" % (kg / 1e9, kd / 1e9, ["%.1e" % e for e in errs], relerr(dv, dv_ref),
                           relerr(ds, ds_ref)), flush=True)
    assert max(errs) <= 1e-12
    assert relerr(dv, dv_ref) <= 1e-10 and relerr(ds, ds_ref) <= 1e-10
    assert kg * 2.4 < kd


@pytest.mark.parametrize("kdtype", ["fp32", "fp16x2"])
def test_ensemble_256_rows_match_oracle(composite, kdtype):
    w, rows = composite
    members, U0 = ensemble_members(w, 256)
    rng = np.random.default_rng(12)
    states = [_perturbed(p, w["ma"], rng) for p in members]
    print("ensemble %s: 256 members, N = 16384" % kdtype, flush=True)
    ens = oq.Ensemble(w["mf"], w["ma"], w["lam"], w["mu"], members, w["pa"], kdtype=kdtype)
    try:
        dU = ens.eval(np.stack([s.pack() for s in states]))
    finally:
        ens.close()
    worst = 0.0
    for m in (0, 1, 63, 64, 127, 128, 200, 255):
        s = states[m]
        dv_ref, ds_ref = rows.rows(members[m], w["pa"], s.x[0], s.x[1], s.x[3])
        got = s.similar().unpack(dU[m])
        dv, ds = rows.pick(got.x[0], got.x[3])
        e = max(relerr(dv, dv_ref), relerr(ds, ds_ref))
        worst = max(worst, e)
        assert e <= 1e-5, "member %d: %g" % (m, e)
    print("ensemble %s rows vs oracle: worst %.2e" % (kdtype, worst), flush=True)


def test_grouped_batch_rows_match_oracle(oracle):
    ws = grouped_batch()
    N = [w["mf"].nx * w["mf"].nxi + (6 * w["ma"].ne if w["ma"] is not None else 0) for w in ws]
    lo, hi = int(np.argmin(N)), int(np.argmax(N))
    rng = np.random.default_rng(13)
    states = [[_perturbed(p, w["ma"], rng) for p in w["members"]] for w in ws]
    print("grouped batch: 24 geometries x 16 members, N %d..%d" % (N[lo], N[hi]), flush=True)
    grp = oq.Group([(w["mf"], w["ma"], w["pa"]) for w in ws], ws[0]["lam"], ws[0]["mu"],
                   [w["members"] for w in ws])
    try:
        dU = grp.unpack(grp.eval(grp.pack([np.stack([s.pack() for s in st]) for st in states])))
    finally:
        grp.close()
    for g in (lo, hi):
        w = ws[g]
        mf, ma = w["mf"], w["ma"]
        I = np.array([0, mf.nx // 2, mf.nx - 1])
        J = np.array([0, mf.nxi // 2, mf.nxi - 1])
        C = [] if ma is None else [0, ma.ne // 2, ma.ne - 1]
        rows = OracleRows(oracle, mf, ma, w["lam"], w["mu"], I, J, C)
        worst = 0.0
        for m in (0, 7, 15):
            s = states[g][m]
            dv_ref, ds_ref = rows.rows(w["members"][m], w["pa"], s.x[0], s.x[1],
                                       s.x[3] if ma is not None else None)
            got = s.similar().unpack(dU[g][m])
            dv, ds = rows.pick(got.x[0], got.x[3] if ma is not None else None)
            e = relerr(dv, dv_ref) if ds is None else max(relerr(dv, dv_ref), relerr(ds, ds_ref))
            worst = max(worst, e)
            assert e <= 1e-5, "geometry %d member %d: %g" % (g, m, e)
        print("grouped geometry %d (N = %d): worst %.2e" % (g, N[g], worst), flush=True)
