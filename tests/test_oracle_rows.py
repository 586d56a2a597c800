"""The sampled-row oracle used by the full-size GPU tests equals the whole
oracle RHS on the same rows (small meshes, CPU)."""
import numpy as np

from helpers import relerr, rs_props, small_fault, small_mantle, visc_props
from oracle_rows import OracleRows

import oetqf_amd as oq


def test_rows_equal_full_coupled_rhs(oracle, rng):
    mf, ma = small_fault(90.0), small_mantle()
    lam = mu = 3e10
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 0.0)
    v = pf.vpl * np.exp(rng.standard_normal((mf.nx, mf.nxi)))
    th = pf.L / v
    sig = 1e6 * rng.standard_normal((ma.ne, 6))
    full = oracle.rhs_coupled(pf, pa, oracle.gf_ff(mf, lam, mu), oracle.gf_mf(mf, ma, lam, mu),
                              oracle.gf_fm(ma, mf, lam, mu), oracle.gf_mm(ma, lam, mu), v, th, sig)
    rows = OracleRows(oracle, mf, ma, lam, mu, [0, 3, 7], [1, 3], [0, 5, ma.ne - 1])
    dv, ds = rows.rows(pf, pa, v, th, sig)
    want_dv, want_ds = rows.pick(full[0], full[3])
    assert relerr(dv, want_dv) <= 1e-13 and relerr(ds, want_ds) <= 1e-13


def test_rows_equal_full_fault_rhs(oracle, rng):
    mf = small_fault(60.0)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(rng.standard_normal((mf.nx, mf.nxi)))
    th = pf.L / v
    full = oracle.rhs_fault(pf, oracle.gf_ff(mf, 3e10, 3e10), v, th)
    rows = OracleRows(oracle, mf, None, 3e10, 3e10, [1, 2, 6], [0, 2])
    dv, _ = rows.rows(pf, None, v, th)
    assert relerr(dv, rows.pick(full[0])[0]) <= 1e-13
