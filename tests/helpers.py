"""Shared builders for the tests (small meshes and properties)."""
import numpy as np

import oetqf_amd as oq


def relerr(a, b):
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = np.asarray(a, dtype=dt)
    b = np.asarray(b, dtype=dt)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def maxrel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def small_fault(dip=90.0, x=20e3, xi=10e3, d=2.5e3):
    return oq.gen_mesh("RectOkada", x, xi, d, d, dip)


def small_mantle(nx=3, ny=2, nz=2):
    return oq.hex8_box(-15e3, -5e3, -12e3, 30e3, 10e3, -20e3, nx, ny, nz)


def rs_props(mf, rng, vpl=1e-9):
    shp = (mf.nx, mf.nxi)
    a = 0.015 + 0.002 * rng.random(shp)
    b = 0.019 + 0.002 * rng.random(shp)
    L = 0.012 * (1 + rng.random(shp))
    sg = 50e6 * (1 + 0.1 * rng.random(shp))
    return oq.RateStateQuasiDynamicProperty(a, b, L, sg, 30e9 / (2 * 3040.0), vpl, 0.6, 1e-6)


def visc_props(ma, rng, n=0.0):
    return oq.PowerLawViscosityProperty(1.0 / (2 * 1e19) * (1 + rng.random(ma.ne)),
                                        n * np.ones(ma.ne), [0.0, -1e-14, 0.0, 0.0, 0.0, 0.0])


def fault_state(mf, p, rng):
    v = p.vpl * np.exp(0.5 * rng.standard_normal((mf.nx, mf.nxi)))
    th = p.L / v * np.exp(0.3 * rng.standard_normal((mf.nx, mf.nxi)))
    dl = rng.random((mf.nx, mf.nxi))
    return v, th, dl


def mantle_state(ma, rng):
    eps = 1e-6 * rng.standard_normal((ma.ne, 6))
    sig = 1e6 * rng.standard_normal((ma.ne, 6))
    return eps, sig
