"""Synthetic earthquake-cycle workloads (meshes, properties, initial state).

Each builder returns a dict with the fault mesh `mf`, optional hex8 mantle
`ma`, elastic constants, friction/viscosity properties and a seeded initial
ArrayPartition.  Randomness uses numpy's PCG64 with the given seed; the
arrays themselves are what tests and benchmarks consume.
"""
import numpy as np

from .equation import ArrayPartition
from .mesh import gen_mesh, gen_mesh_graded, hex8_box
from .property import PowerLawViscosityProperty, RateStateQuasiDynamicProperty

YEAR = 365 * 86400.0


def _b_profile(depth_m, top=2e3, bottom=12e3, taper=2e3, b_vw=0.019, b_vs=0.011):
    """b = b_vw between `top` and `bottom`, tapering linearly to b_vs over
    `taper` above `top` and below `bottom`, b_vs elsewhere."""
    d = np.asarray(depth_m, dtype=np.float64)
    b = np.full(d.shape, b_vs)
    up = d < top
    b[up] = b_vs + (b_vw - b_vs) * d[up] / top
    mid = (d >= top) & (d <= bottom)
    b[mid] = b_vw
    lo = (d > bottom) & (d < bottom + taper)
    b[lo] = b_vw - (b_vw - b_vs) * (d[lo] - bottom) / taper
    return b


def strike_slip_512(seed=20261202):
    """Vertical strike-slip fault, 40 km x 20 km, 32 x 16 = 512 Okada patches
    in an elastic half-space; mu = 30 GPa, nu = 0.25, cs = 3.04 km/s,
    eta = mu/(2 cs), vpl = 1e-9 m/s, sigma = 50 MPa, L = 12 mm, a = 0.015,
    b = 0.019 at 2-12 km tapering to 0.011; V0 = vpl exp(0.1 N(0,1)),
    theta0 = L/V0.  Tsit5 reltol 1e-6 (abstol 1e-8), 300 years."""
    rng = np.random.default_rng(seed)
    mf = gen_mesh("RectOkada", 40e3, 20e3, 1250.0, 1250.0, 90.0)
    mu = 30e9
    lam = mu  # nu = 0.25
    shp = (mf.nx, mf.nxi)
    depth = np.broadcast_to(-mf.z, shp)
    a = np.full(shp, 0.015)
    b = _b_profile(depth)
    pf = RateStateQuasiDynamicProperty(a, b, np.full(shp, 0.012), np.full(shp, 50e6),
                                       mu / (2 * 3040.0), 1e-9, 0.6, 1e-6)
    v0 = pf.vpl * np.exp(0.1 * rng.standard_normal(shp))
    u0 = ArrayPartition(v0, pf.L / v0, np.zeros(shp))
    return dict(name="strike-slip fault, 512 patches (32x16), elastic", mf=mf, ma=None, lam=lam,
                mu=mu, pf=pf, pa=None, u0=u0, tspan=(0.0, 300 * YEAR), nrept=2, buffer_ratio=0.0,
                reltol=1e-6, abstol=1e-8)


def transform_fault_composite(nx=128, nxi=32, mx=16, my=16, mz=8, seed=20261202,
                              grading=1.05, length=80e3, depth=20e3):
    """Transform fault `length` x `depth` on nx x nxi patches (down-dip cells
    graded finer toward the surface by `grading`) coupled to a viscoelastic
    mantle block below 10 km (80 x 40 x 30 km, mx x my x mz hex8 cells, six
    strain components each), linear Maxwell viscosity 1e19 Pa s.  Friction
    as strike_slip_512 with velocity strengthening below 10 km.  Defaults:
    4096 patches + 2048 cells (N = 16384)."""
    rng = np.random.default_rng(seed)
    mf = gen_mesh_graded(length, depth, length / nx, nxi, 90.0, ratio=grading)
    ma = hex8_box(-40e3, -20e3, -10e3, 80e3, 40e3, -30e3, mx, my, mz)
    mu = 30e9
    lam = mu
    shp = (mf.nx, mf.nxi)
    d = np.broadcast_to(-mf.z, shp)
    pf = RateStateQuasiDynamicProperty(np.full(shp, 0.015), _b_profile(d, bottom=8e3),
                                       np.full(shp, 0.012), np.full(shp, 50e6),
                                       mu / (2 * 3040.0), 1e-9, 0.6, 1e-6)
    pa = PowerLawViscosityProperty(np.full(ma.ne, 1.0 / (2 * 1e19)), np.zeros(ma.ne), np.zeros(6))
    v0 = pf.vpl * np.exp(0.1 * rng.standard_normal(shp))
    u0 = ArrayPartition(v0, pf.L / v0, np.zeros((ma.ne, 6)), np.zeros((ma.ne, 6)), np.zeros(shp))
    return dict(name="transform fault %dx%d=%d patches + hex8 mantle %dx%dx%d=%d cells"
                % (nx, nxi, nx * nxi, mx, my, mz, ma.ne),
                mf=mf, ma=ma, lam=lam, mu=mu, pf=pf, pa=pa, u0=u0, tspan=(0.0, 500 * YEAR),
                nrept=2, buffer_ratio=0.0, reltol=1e-6, abstol=1e-8)


def ensemble_members(w, M=256, seed=20261203):
    """M frictional realizations on workload w's geometry: in the
    seismogenic zone (where the base b exceeds a) a - b ~ U[-0.006, -0.002];
    L ~ logU[4, 40] mm; sigma ~ U[30, 80] MPa (per-member scalars); initial
    V = vpl exp(0.1 N(0,1)) per patch, theta = L/V.  Returns (members, U0)
    with U0 (M, nstate) packed ArrayPartition states."""
    rng = np.random.default_rng(seed)
    base = w["pf"]
    shp = base.a.shape
    vw = base.b > base.a
    members, states = [], []
    for _ in range(M):
        amb = rng.uniform(-0.006, -0.002)
        L = np.exp(rng.uniform(np.log(4e-3), np.log(4e-2)))
        sg = rng.uniform(30e6, 80e6)
        b = base.b.copy()
        b[vw] = base.a[vw] - amb
        p = RateStateQuasiDynamicProperty(base.a.copy(), b, np.full(shp, L), np.full(shp, sg),
                                          base.η, base.vpl, base.f0, base.v0)
        v0 = p.vpl * np.exp(0.1 * rng.standard_normal(shp))
        parts = [v0, p.L / v0]
        if w["ma"] is not None:
            ne = w["ma"].ne
            parts += [np.zeros((ne, 6)), np.zeros((ne, 6))]
        parts.append(np.zeros(shp))
        members.append(p)
        states.append(ArrayPartition(*parts).pack())
    return members, np.stack(states)


def precision_scale_65536(seed=20261202):
    """The composite model at scale: 256 x 64 = 16384 fault patches coupled
    to a 32 x 32 x 8 = 8192-cell mantle (49152 strain dofs), N = 65536; the
    dense operator is 34.4 GB in FP64.  Physics as transform_fault_composite;
    200 years."""
    w = transform_fault_composite(nx=256, nxi=64, mx=32, my=32, mz=8, seed=seed)
    w["tspan"] = (0.0, 200 * YEAR)
    return w


def grouped_batch(G=24, M=16, seed=20261204, pmin=1024, pmax=8192):
    """G synthetic faults of uneven size: patch counts log-uniform in
    [pmin, pmax] at aspect ratio 4:1 (80 x 20 km; nxi = round(sqrt(P/4)),
    nx = 4 nxi), planar (uniform rows) and transfinite (rows graded x1.05)
    alternating; half of them, drawn at random, coupled to a hex8 mantle
    block below 10 km (80 x 40 x 30 km) of 512-4096 cells (log-uniform,
    m x m x 8), linear Maxwell viscosity 1e19 Pa s.  M frictional
    realizations per geometry drawn as ensemble_members.  Returns a list of
    workload dicts (each with "members" and "U0")."""
    rng = np.random.default_rng(seed)
    coupled = set(rng.permutation(G)[:G // 2].tolist())
    out = []
    for g in range(G):
        P = float(np.exp(rng.uniform(np.log(pmin), np.log(pmax))))
        nxi = max(4, int(round(np.sqrt(P / 4.0))))
        nx = 4 * nxi
        grading = 1.0 if g % 2 == 0 else 1.05
        mf = gen_mesh_graded(80e3, 20e3, 80e3 / nx, nxi, 90.0, ratio=grading)
        mu = 30e9
        shp = (mf.nx, mf.nxi)
        d = np.broadcast_to(-mf.z, shp)
        ma = pa = None
        if g in coupled:
            ne = float(np.exp(rng.uniform(np.log(512), np.log(4096))))
            m = max(2, int(round(np.sqrt(ne / 8.0))))
            ma = hex8_box(-40e3, -20e3, -10e3, 80e3, 40e3, -30e3, m, m, 8)
            pa = PowerLawViscosityProperty(np.full(ma.ne, 1.0 / (2 * 1e19)), np.zeros(ma.ne),
                                           np.zeros(6))
        b = _b_profile(d, bottom=8e3) if ma is not None else _b_profile(d)
        pf = RateStateQuasiDynamicProperty(np.full(shp, 0.015), b, np.full(shp, 0.012),
                                           np.full(shp, 50e6), mu / (2 * 3040.0), 1e-9, 0.6, 1e-6)
        parts = [np.zeros(shp), np.zeros(shp)]
        if ma is not None:
            parts += [np.zeros((ma.ne, 6)), np.zeros((ma.ne, 6))]
        parts.append(np.zeros(shp))
        w = dict(name="fault %dx%d%s" % (nx, nxi, " + mantle %d cells" % ma.ne if ma is not None else ""),
                 mf=mf, ma=ma, lam=mu, mu=mu, pf=pf, pa=pa, u0=ArrayPartition(*parts),
                 tspan=(0.0, 300 * YEAR), nrept=2, buffer_ratio=0.0, reltol=1e-6, abstol=1e-8)
        w["members"], w["U0"] = ensemble_members(w, M, seed=seed + 1 + g)
        out.append(w)
    return out
