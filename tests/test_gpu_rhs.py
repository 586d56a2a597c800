"""GPU RHS `f(du,u,p,t)` vs the CPU oracle (FP64 <= 1e-12 relative).

Reference: src/BEM/equation.jl:156-166 (fault only), :185-205 (coupled);
the GPU path is the fused prologue + K-streaming kernel behind the C ABI.
"""
import numpy as np
import pytest

import oetqf_amd as oq
from helpers import fault_state, mantle_state, relerr, rs_props, small_fault, small_mantle, visc_props

pytestmark = pytest.mark.gpu


def _coupled_inputs(oracle, rng, dip=90.0, n=0.0):
    mf, ma = small_fault(dip), small_mantle()
    lam = mu = 3e10
    st = oracle.gf_ff(mf, lam, mu, buffer_ratio=1.0)
    g12 = oracle.gf_mf(mf, ma, lam, mu, buffer_ratio=1.0)
    g21 = oracle.gf_fm(ma, mf, lam, mu)
    g22 = oracle.gf_mm(ma, lam, mu)
    pf, pa = rs_props(mf, rng), visc_props(ma, rng, n)
    v, th, dl = fault_state(mf, pf, rng)
    eps, sig = mantle_state(ma, rng)
    return mf, ma, st, g12, g21, g22, pf, pa, (v, th, eps, sig, dl)


def _oracle_coupled(oracle, pf, pa, st, g12, g21, g22, u):
    v, th, eps, sig, dl = u
    return oracle.rhs_coupled(pf, pa, st, g12, g21, g22, v, th, sig)


@pytest.mark.parametrize("fourier", [False, True])
def test_rhs_fault_only(oq, oracle, rng, fourier):
    mf = small_fault(90.0)
    st = oracle.gf_ff(mf, 3e10, 3e10, buffer_ratio=1.0)
    gf = np.fft.rfft(np.concatenate([st, st[:0:-1]], axis=0), axis=0) if fourier else st
    pf = rs_props(mf, rng)
    v, th, dl = fault_state(mf, pf, rng)
    u0 = oq.ArrayPartition(v, th, dl)
    prob = oq.assemble(gf, pf, u0, (0.0, 1.0))
    du = oq.similar(u0)
    prob.f(du, u0, prob.p, 1.0)
    dv, dth, dd = oracle.rhs_fault(pf, st, v, th)
    assert relerr(du.x[0], dv) <= 1e-12
    assert relerr(du.x[1], dth) <= 1e-14
    assert np.array_equal(du.x[2], v)


@pytest.mark.parametrize("n", [0.0, 2.5])
def test_rhs_coupled(oq, oracle, rng, n):
    mf, ma, st, g12, g21, g22, pf, pa, u = _coupled_inputs(oracle, rng, n=n)
    u0 = oq.ArrayPartition(*u)
    prob = oq.assemble(st, g12, g21, g22, pf, pa, u0, (0.0, 1.0))
    du = oq.similar(u0)
    prob.f(du, u0, prob.p, 1.0)
    ref = _oracle_coupled(oracle, pf, pa, st, g12, g21, g22, u)
    for k, (got, want) in enumerate(zip((du.x[0], du.x[1], du.x[2], du.x[3], du.x[4]), ref)):
        assert relerr(got, want) <= 1e-12, "component %d" % k


def test_rhs_composite_viscosity(oq, oracle, rng):
    mf, ma, st, g12, g21, g22, pf, pa, u = _coupled_inputs(oracle, rng, n=0.0)
    p2 = oq.PowerLawViscosityProperty(1e-30 * np.ones(ma.ne), 2.0 * np.ones(ma.ne), pa.dϵ0)
    pc = oq.CompositePowerLawViscosityProperty([pa, p2], pa.dϵ0)
    u0 = oq.ArrayPartition(*u)
    prob = oq.assemble(st, g12, g21, g22, pf, pc, u0, (0.0, 1.0))
    du = oq.similar(u0)
    prob.f(du, u0, prob.p, 1.0)
    sig = u[3]
    e1 = oracle.update_strain_rate(pa, sig)
    e2 = oracle.update_strain_rate(p2, sig)
    assert relerr(du.x[2], e1 + e2) <= 1e-14


def test_device_assembly_matches_host_blocks(oq, oracle, rng):
    """Green's functions built straight into the device operator equal the
    host-assembled operator, and both RHS agree."""
    mf, ma, st, g12, g21, g22, pf, pa, u = _coupled_inputs(oracle, rng)
    u0 = oq.ArrayPartition(*u)
    p1 = oq.assemble(st, g12, g21, g22, pf, pa, u0, (0.0, 1.0))
    p2 = oq.assemble_from_mesh(mf, ma, 3e10, 3e10, pf, pa, u0, (0.0, 1.0), buffer_ratio=1.0)
    k1, k2 = p1.f.dense_operator(), p2.f.dense_operator()
    assert relerr(k2, k1) <= 1e-10
    du1, du2 = oq.similar(u0), oq.similar(u0)
    p1.f(du1, u0, None, 0.0)
    p2.f(du2, u0, None, 0.0)
    for a, b in zip(du1.x, du2.x):
        assert relerr(a, b) <= 1e-9


def test_dense_operator_layout(oq, oracle, rng):
    """K[(i,j),(k,l)] = st[|i-k|, j, l] and the mantle blocks in receiver-
    major order (test/BEM/tests.jl:46-49; GF.jl:163-169,212,268)."""
    mf, ma, st, g12, g21, g22, pf, pa, u = _coupled_inputs(oracle, rng)
    prob = oq.assemble(st, g12, g21, g22, pf, pa, oq.ArrayPartition(*u), (0.0, 1.0))
    K = prob.f.dense_operator()
    nx, nxi = mf.nx, mf.nxi
    nf = nx * nxi
    i, j, k, l = np.meshgrid(np.arange(nx), np.arange(nxi), np.arange(nx), np.arange(nxi), indexing="ij")
    kff = st[np.abs(i - k), j, l].transpose(1, 0, 3, 2).reshape(nf, nf)
    assert np.array_equal(K[:nf, :nf], kff)
    assert np.array_equal(K[:nf, nf:], g21)
    assert np.array_equal(K[nf:, :nf], g12)
    assert np.array_equal(K[nf:, nf:], g22)


def _round_nearest_even(x, p):
    """x rounded to p significant bits, ties to even (normal range): with
    |x| = m 2^e, m in [0.5, 1), the candidates are floor(m 2^p) and the next
    integer, scaled back; the distances are exact in FP64."""
    m, e = np.frexp(np.abs(x))
    q = m * 2.0 ** p
    lo = np.floor(q)
    f = q - lo
    up = (f > 0.5) | ((f == 0.5) & (np.fmod(lo, 2.0) == 1.0))
    return np.sign(x) * np.ldexp((lo + up) / 2.0 ** p, e)


@pytest.mark.parametrize("kdtype,tol,exact_tol", [("fp32", 1e-5, 1e-12), ("bf16", 5e-2, 1e-6)])
def test_rhs_reduced_precision_operator(oq, oracle, rng, kdtype, tol, exact_tol):
    """K streamed as FP32/BF16 with FP64 row accumulation: the RHS error
    against FP64 is bounded by the storage rounding of K (tol, per dtype).
    Against the FP64 oracle applied to the rounded operator the FP32 path
    agrees to FP64 accuracy; BF16 forms its products in FP32 against an
    FP32 copy of the source vector (rounding 2^-24), hence exact_tol."""
    mf, ma, st, g12, g21, g22, pf, pa, u = _coupled_inputs(oracle, rng)
    u0 = oq.ArrayPartition(*u)
    p64 = oq.assemble(st, g12, g21, g22, pf, pa, u0, (0.0, 1.0))
    plo = oq.assemble(st, g12, g21, g22, pf, pa, u0, (0.0, 1.0), kdtype=kdtype)
    d64, dlo = oq.similar(u0), oq.similar(u0)
    p64.f(d64, u0, None, 0.0)
    plo.f(dlo, u0, None, 0.0)
    # exact check: the low-precision RHS equals the FP64 oracle applied to the
    # rounded operator, to FP64 accuracy
    K = plo.f.dense_operator()
    K64 = p64.f.dense_operator()
    # the stored operator is the FP64 one rounded to nearest even
    assert np.array_equal(K, _round_nearest_even(K64, 24 if kdtype == "fp32" else 8))
    nf = mf.nx * mf.nxi
    v, th, eps, sig, dl = u
    ref = oracle.rhs_dense(pf, pa, K, nf, ma.ne, v.reshape(-1, order="F"), th.reshape(-1, order="F"),
                           sig.reshape(-1, order="F"))
    assert relerr(dlo.x[0].reshape(-1, order="F"), ref[0]) <= exact_tol
    assert relerr(dlo.x[3].reshape(-1, order="F"), ref[3]) <= exact_tol
    err = relerr(dlo.x[3], d64.x[3])
    print("%s mantle stress-rate rel err vs FP64: %.3e" % (kdtype, err))
    assert err <= tol


def test_rhs_nonpositive_state_variable(oq, oracle, rng):
    """θ <= 0: ψ1 = exp(-Inf) = 0 (max(0,θ), equation.jl:238) so dv = dτ/η."""
    mf = small_fault(90.0)
    st = oracle.gf_ff(mf, 3e10, 3e10)
    pf = rs_props(mf, rng)
    v, th, dl = fault_state(mf, pf, rng)
    th[0, 0] = -1.0
    u0 = oq.ArrayPartition(v, th, dl)
    prob = oq.assemble(st, pf, u0, (0.0, 1.0))
    du = oq.similar(u0)
    prob.f(du, u0, None, 0.0)
    dv, dth, _ = oracle.rhs_fault(pf, st, v, th)
    assert relerr(du.x[0], dv) <= 1e-12
    s = (v - pf.vpl).reshape(-1, order="F")
    dtau = oracle.dtau_toeplitz(st, (v - pf.vpl))
    assert abs(du.x[0][0, 0] - dtau[0, 0] / pf.η) <= 1e-12 * abs(dtau[0, 0] / pf.η)
    assert s.size == v.size


def test_rhs_argument_errors(oq, oracle, rng):
    mf = small_fault(90.0)
    st = oracle.gf_ff(mf, 3e10, 3e10)
    pf = rs_props(mf, rng)
    v, th, dl = fault_state(mf, pf, rng)
    with pytest.raises(AssertionError):
        oq.RateStateQuasiDynamicProperty(pf.a, pf.b, pf.L, pf.σ, -1.0, pf.vpl)
    with pytest.raises(ValueError):
        oq.assemble(st, pf, oq.ArrayPartition(v, th, dl), (0.0, 1.0), kdtype="fp8")
