"""Fault-only problems with K_ff kept as its Toeplitz kernel, applied
directly (OQ_K_TOEPLITZ) or by DFT convolution along strike as the
reference applies it (OQ_K_FFT, equation.jl:44-61): the RHS against the
oracle's DFT convolution and against the dense FP64 operator; the operator
it expands to is the dense one; trajectories agree with the dense
stepper."""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import relerr, rs_props, small_fault
from oetqf_amd.workloads import YEAR

pytestmark = pytest.mark.gpu


def _state(mf, pf, rng):
    v = pf.vpl * np.exp(rng.standard_normal((mf.nx, mf.nxi)))
    th = pf.L / v * np.exp(0.3 * rng.standard_normal((mf.nx, mf.nxi)))
    return oq.ArrayPartition(v, th, rng.random((mf.nx, mf.nxi)))


@pytest.mark.parametrize("kdtype", ["toeplitz", "fft"])
@pytest.mark.parametrize("dip,fourier", [(90.0, False), (45.0, True)])
def test_toeplitz_rhs_matches_oracle_fft(oracle, rng, dip, fourier, kdtype):
    mf = small_fault(dip, x=40e3, xi=15e3, d=1.25e3)
    pf = rs_props(mf, rng)
    u = _state(mf, pf, rng)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    gf = oq.stress_greens_function(mf, 3e10, 3e10, buffer_ratio=1.0, fourier=fourier)
    prob = oq.assemble(gf, pf, u, (0.0, 1.0), kdtype=kdtype)
    du = oq.similar(u)
    prob.f(du, u, None, 0.0)
    ref = oracle.rhs_fault(pf, st, u.x[0], u.x[1])
    for k in range(3):
        assert relerr(du.x[k], ref[k]) <= 1e-12, k
    prob.f.close()


@pytest.mark.parametrize("kdtype", ["toeplitz", "fft"])
def test_toeplitz_from_mesh_equals_dense(rng, kdtype):
    mf = oq.gen_mesh("RectOkada", 80e3, 20e3, 312.5, 312.5, 90.0)  # 256 x 64 = 16384 patches
    pf = rs_props(mf, rng)
    u = _state(mf, pf, rng)
    pt = oq.assemble_from_mesh(mf, None, 3e10, 3e10, pf, None, u, (0.0, 1.0), kdtype=kdtype)
    pd = oq.assemble_from_mesh(mf, None, 3e10, 3e10, pf, None, u, (0.0, 1.0))
    dt_, dd = oq.similar(u), oq.similar(u)
    pt.f(dt_, u, None, 0.0)
    pd.f(dd, u, None, 0.0)
    e = relerr(dt_.x[0], dd.x[0])
    mt, md = pt.f.time(20)[0], pd.f.time(20)[0]
    print("16384 patches: %s %.4f ms, dense FP64 %.4f ms per RHS, rel diff %.2e" % (kdtype, mt, md, e))
    assert e <= 1e-12
    pt.f.close()
    pd.f.close()


@pytest.mark.parametrize("kdtype", ["toeplitz", "fft"])
def test_toeplitz_operator_and_trajectory(oracle, rng, kdtype):
    mf = small_fault(90.0)
    pf = rs_props(mf, rng)
    v = pf.vpl * np.exp(0.3 * rng.standard_normal((mf.nx, mf.nxi)))
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros_like(v))
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    pt = oq.assemble(st, pf, u0, (0.0, 0.5 * YEAR), kdtype=kdtype)
    pd = oq.assemble(st, pf, u0, (0.0, 0.5 * YEAR))
    assert np.array_equal(pt.f.dense_operator(), pd.f.dense_operator())
    st_ = oq.solve(pt, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    sd = oq.solve(pd, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    assert st_.retcode == sd.retcode == "Success"
    assert abs(st_.naccept - sd.naccept) <= 1
    assert relerr(st_.u.pack(), sd.u.pack()) <= 1e-8


def test_toeplitz_rejects_coupled(oq, oracle, rng):
    """The direct Toeplitz operator is fault-only; a mantle is refused."""
    from helpers import small_mantle, visc_props
    mf, ma = small_fault(90.0), small_mantle()
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 0.0)
    v = pf.vpl * np.ones((mf.nx, mf.nxi))
    u0 = oq.ArrayPartition(v, pf.L / v, np.zeros((ma.ne, 6)), np.zeros((ma.ne, 6)), np.zeros_like(v))
    with pytest.raises(ValueError):
        oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u0, (0.0, 1.0), kdtype="toeplitz")


def _coupled_state(mf, ma, pf, rng):
    v, th, dl = (pf.vpl * np.exp(rng.standard_normal((mf.nx, mf.nxi))),
                 pf.L / pf.vpl * np.exp(0.3 * rng.standard_normal((mf.nx, mf.nxi))),
                 rng.random((mf.nx, mf.nxi)))
    return oq.ArrayPartition(v, th, np.zeros((ma.ne, 6)), 1e6 * rng.standard_normal((ma.ne, 6)), dl)


@pytest.mark.parametrize("fourier", [False, True])
def test_fft_coupled_rhs_matches_oracle(oracle, rng, fourier):
    """The coupled model with K_ff applied by DFT convolution, as the
    reference's coupled ode applies it (equation.jl:185-205, dτ_dt! by FFT
    at :200), the other three blocks dense: every component of du against
    the oracle's coupled RHS, and against the same model with K dense."""
    from helpers import small_mantle, visc_props
    mf, ma = small_fault(60.0, x=40e3, xi=15e3, d=1.25e3), small_mantle(4, 2, 3)
    lam = mu = 3e10
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 3.0)
    u = _coupled_state(mf, ma, pf, rng)
    st = oracle.gf_ff(mf, lam, mu, buffer_ratio=1.0)
    gf11 = oq.stress_greens_function(mf, lam, mu, buffer_ratio=1.0, fourier=fourier)
    g12, g21, g22 = (oracle.gf_mf(mf, ma, lam, mu, buffer_ratio=1.0), oracle.gf_fm(ma, mf, lam, mu),
                     oracle.gf_mm(ma, lam, mu))
    pt = oq.assemble(gf11, g12, g21, g22, pf, pa, u, (0.0, 1.0), kdtype="fft")
    pd = oq.assemble(gf11, g12, g21, g22, pf, pa, u, (0.0, 1.0))  # the same K_ff input, dense
    dt_, dd = oq.similar(u), oq.similar(u)
    pt.f(dt_, u, None, 0.0)
    pd.f(dd, u, None, 0.0)
    ref = oracle.rhs_coupled(pf, pa, st, g12, g21, g22, u.x[0], u.x[1], u.x[3])
    for k in range(5):
        assert relerr(dt_.x[k], ref[k]) <= 1e-12, k
        assert relerr(dt_.x[k], dd.x[k]) <= 1e-12, k
    assert np.array_equal(pt.f.dense_operator(), pd.f.dense_operator())
    pt.f.close()
    pd.f.close()


def test_fft_coupled_from_mesh_and_trajectory(rng):
    """Device assembly of the coupled DFT model equals the dense device
    assembly (operator bit for bit, RHS to 1e-12), and a Tsit5 trajectory
    on it follows the dense one."""
    from helpers import small_mantle, visc_props
    mf, ma = small_fault(90.0), small_mantle()
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 0.0)
    u = _coupled_state(mf, ma, pf, rng)
    kw = dict(buffer_ratio=1.0, qtype="Gauss2")
    pt = oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u, (0.0, 0.05 * YEAR), kdtype="fft", **kw)
    pd = oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u, (0.0, 0.05 * YEAR), **kw)
    assert np.array_equal(pt.f.dense_operator(), pd.f.dense_operator())
    dt_, dd = oq.similar(u), oq.similar(u)
    pt.f(dt_, u, None, 0.0)
    pd.f(dd, u, None, 0.0)
    for k in range(5):
        assert relerr(dt_.x[k], dd.x[k]) <= 1e-12, k
    st_ = oq.solve(pt, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    sd = oq.solve(pd, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    assert st_.retcode == sd.retcode == "Success"
    assert abs(st_.naccept - sd.naccept) <= 1
    assert relerr(st_.u.pack(), sd.u.pack()) <= 1e-8
    pt.f.close()
    pd.f.close()


def _blocks(K, nf):
    return K[:nf], K[nf:, :nf], K[nf:, nf:]


@pytest.mark.parametrize("dims,rfzh", [((5, 3, 3), [1.0, 1.5, 2.25]), ((4, 4, 2), None)])
def test_fft_grid_matches_dense_and_oracle(oracle, rng, dims, rfzh):
    """OQ_K_FFT_GRID: K_ff by DFT, K_mm as 2D convolutions of its per-offset
    kernel over a mantle on a uniform horizontal grid (graded layers
    allowed).  With 5 cells over 40 km the mantle spacing (8 km) is 6.4
    fault spacings and K_fm, K_mf stay dense (bit for bit the dense
    assembly's); with 4 cells (10 km = 8 spacings) they are kept as tables
    along strike.  Table-built blocks equal the dense assembly within 1e-12
    (shifted virtual receivers, equal to rounding); every RHS component is
    within 1e-12 of the dense model's and of the oracle's coupled RHS on the
    oracle's blocks (within the two assemblies' 1e-10 agreement), and a
    Tsit5 trajectory follows the dense one."""
    from helpers import visc_props
    mf = small_fault(60.0, x=40e3, xi=15e3, d=1.25e3)
    ma = oq.hex8_box(-20e3, -6e3, -10e3, 40e3, 12e3, -16e3, *dims, rfzh=rfzh)
    lam = mu = 3e10
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, 0.0)
    u = _coupled_state(mf, ma, pf, rng)
    kw = dict(buffer_ratio=1.0)
    pg = oq.assemble_from_mesh(mf, ma, lam, mu, pf, pa, u, (0.0, 0.05 * YEAR), kdtype="fft_grid", **kw)
    pd = oq.assemble_from_mesh(mf, ma, lam, mu, pf, pa, u, (0.0, 0.05 * YEAR), **kw)
    nf = mf.nx * mf.nxi
    gk, dk = _blocks(pg.f.dense_operator(), nf), _blocks(pd.f.dense_operator(), nf)
    assert np.array_equal(gk[0][:, :nf], dk[0][:, :nf])  # K_ff
    close = lambda a, b: np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    if dims[0] == 5:
        assert np.array_equal(gk[0], dk[0]) and np.array_equal(gk[1], dk[1])
    else:
        assert close(gk[0][:, nf:], dk[0][:, nf:]) and close(gk[1], dk[1])
    assert close(gk[2], dk[2])
    _, _, kb = pg.f.info()
    assert kb < pd.f.info()[2]
    dg, dd = oq.similar(u), oq.similar(u)
    pg.f(dg, u, None, 0.0)
    pd.f(dd, u, None, 0.0)
    st = oracle.gf_ff(mf, lam, mu, buffer_ratio=1.0)
    g12, g21, g22 = (oracle.gf_mf(mf, ma, lam, mu, buffer_ratio=1.0), oracle.gf_fm(ma, mf, lam, mu),
                     oracle.gf_mm(ma, lam, mu))
    ref = oracle.rhs_coupled(pf, pa, st, g12, g21, g22, u.x[0], u.x[1], u.x[3])
    for k in range(5):
        assert relerr(dg.x[k], dd.x[k]) <= 1e-12, k
        assert relerr(dg.x[k], ref[k]) <= 1e-10, k
    sg = oq.solve(pg, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    sd = oq.solve(pd, oq.Tsit5(), reltol=1e-6, abstol=1e-8)
    assert sg.retcode == sd.retcode == "Success"
    assert abs(sg.naccept - sd.naccept) <= 1
    assert relerr(sg.u.pack(), sd.u.pack()) <= 1e-8
    pg.f.close()
    pd.f.close()
