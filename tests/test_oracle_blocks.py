"""Pins of the oracle's four Green's block builders (oracle/greens.c:
oq_gf_ff, oq_gf_mf, oq_gf_fm, oq_gf_mm_rows), which restate
src/BEM/GF.jl:31-58 (K_ff), :123-174 (K_mf = gf₁₂), :194-227 (K_fm = gf₂₁)
and :250-296 (K_mm = gf₂₂).

The kernels underneath (Okada's dc3d and the hex8 strain volume) are pinned
in test_oracle_okada.py / test_oracle_hex8.py.  Here the assembly on top of
them -- traction projections, periodic images, quadrature mapping, block and
index layout -- is checked against what the elasticity fixes, computed
independently of the builders' component algebra:

- additivity: summing K_ff (K_mf) along strike must equal one long
  rectangle (or the 2P+1 separated ones with a buffer), which fixes the image
  period lrept, the image range, the source column and the row indices;
- the 2-D antiplane limit: a long vertical strike-slip fault summed along
  strike is a strip of screw dislocations with its free-surface image;
- stress drop: a slipping patch unloads itself (negative K_ff diagonal), for
  strike and dip slip at every dip;
- constitutive law: K_mf rows are λ tr(ε) I + 2μ ε of the FD-pinned
  displacement gradients, in the order xx xy xz yy yz zz;
- quadrature: K_mf / K_mm with a Gauss rule equal the cell average taken by a
  midpoint rule on subcells (Richardson-extrapolated);
- Betti reciprocity: the work of the fault's stress on a cell eigenstrain
  equals the work of the cell's traction on the fault slip, which ties K_mf
  to K_fm (and K_mm to itself) including every sign and the engineering
  factor 2 of shear components;
- the slab sum rule: a laterally extended layer of cells with one uniform
  eigenstrain gives -C:ε* relaxed to σ_i3 = 0 inside and nothing outside;
- the reference's stability diagnostic (GF.jl:291-294): with cell-averaged
  rows max Re eig(K_mm) < 0 (an energy argument).

Each builder has planted mistakes (a flipped sign, a transposed index, a
wrong offset or period) in a rebuilt copy of oracle/greens.c; every one of
them must fail at least one pin (test_planted_mistakes_are_caught).
"""
import ctypes
import os
import shutil
import subprocess
from contextlib import contextmanager
from types import SimpleNamespace

import numpy as np
import pytest

import oetqf_amd as oq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LAM, MU = 3e10, 3e10
ALPHA = (LAM + MU) / (LAM + 2 * MU)
SS, DS = 0, 1
PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]  # xx xy xz yy yz zz (GF.jl:164-169)
W_ENG = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])         # σ:E_p = W_ENG[p] σ_p
_HEX = ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")


# ---------------- independent elasticity helpers ----------------

def grad_matrix(u):
    """G[a, b] = du_a/dx_b from dc3d's 12-vector (layout pinned by finite
    differences in test_oracle_okada.py)."""
    return np.array([[u[3], u[6], u[9]], [u[4], u[7], u[10]], [u[5], u[8], u[11]]])


def stress_tensor(u):
    e = 0.5 * (grad_matrix(u) + grad_matrix(u).T)
    return LAM * np.trace(e) * np.eye(3) + 2 * MU * e


def vec6(s):
    return np.array([s[a, b] for a, b in PAIRS])


def tensor6(v):
    s = np.zeros((3, 3))
    for (a, b), x in zip(PAIRS, v):
        s[a, b] = s[b, a] = x
    return s


def frame(dip):
    """Okada's fault frame: normal n, strike and down-dip slip directions
    (the displacement jumps by +b across the plane along n,
    test_unit_slip_jump_sign)."""
    sd, cd = np.sin(np.radians(dip)), np.cos(np.radians(dip))
    return np.array([0.0, -sd, cd]), np.array([1.0, 0.0, 0.0]), np.array([0.0, cd, sd])


def shear_traction(s, ftype, dip):
    n, es, ed = frame(dip)
    return (es if ftype == SS else ed) @ s @ n


def slip(ftype):
    return (1.0, 0.0, 0.0) if ftype == SS else (0.0, 1.0, 0.0)


def dc3d(O, x, y, z, dip, ftype, al, aw):
    u, ret = O.dc3d(x, y, z, ALPHA, 0.0, dip, al[0], al[1], aw[0], aw[1], *slip(ftype))
    assert ret == 0
    return u


def cells(spec):
    """Hex8 cells from (cx, cy, cz, dx, dy, dz) rows; q = (cx, cy - dy/2,
    cz + dz/2) (mesh.jl:181-183)."""
    a = np.asarray(spec, dtype=np.float64)
    c = dict(cx=a[:, 0], cy=a[:, 1], cz=a[:, 2], dx=a[:, 3], dy=a[:, 4], dz=a[:, 5])
    c.update(qx=c["cx"], qy=c["cy"] - c["dy"] / 2, qz=c["cz"] + c["dz"] / 2)
    return SimpleNamespace(**c, ne=len(a))


def refine(mf, m):
    """The fault mesh with every patch split into m x m equal sub-patches."""
    k = np.arange(m)
    ax = np.concatenate([[(a + (b - a) * i / m, a + (b - a) * (i + 1) / m) for i in k] for a, b in mf.ax])
    axi = np.concatenate([[(a + (b - a) * i / m, a + (b - a) * (i + 1) / m) for i in k] for a, b in mf.axi])
    xi = axi.mean(axis=1)
    sd, cd = np.sin(np.radians(mf.dip)), np.cos(np.radians(mf.dip))
    return SimpleNamespace(x=ax.mean(axis=1), ax=ax, y=xi * cd, z=xi * sd, axi=axi, nx=mf.nx * m,
                           nxi=mf.nxi * m, dx=mf.dx / m, dep=0.0, dip=mf.dip)


def midpoint_average(f, c, i, n):
    """Average of f(x, y, z) (a 6-vector) over cell i by the n^3 midpoint rule."""
    t = (np.arange(n) + 0.5) / n - 0.5
    acc = np.zeros(6)
    for a in t:
        for b in t:
            for g in t:
                acc += f(c.cx[i] + a * c.dx[i], c.cy[i] + b * c.dy[i], c.cz[i] + g * c.dz[i])
    return acc / n ** 3


def richardson_average(f, c, i, n):
    return (4 * midpoint_average(f, c, i, 2 * n) - midpoint_average(f, c, i, n)) / 3


def hex8_stress(O, x, y, z, c, j, p):
    e = np.zeros(6)
    e[p] = 1.0
    return O.stress_vol_hex8(x, y, z, c.qx[j], c.qy[j], c.qz[j], c.dx[j], c.dy[j], c.dz[j], e, MU,
                             LAM / 2 / (LAM + MU))


def relnorm(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


# ---------------- pins of K_ff (GF.jl:31-58) ----------------

def pin_ff_strike_sum(O, ftype, dip, beta, nrept=2):
    """sum_i st[i, j, l]: by translation invariance, the traction at x_1
    from the 2P+1 images of a strike run of nx columns -- one rectangle when
    the buffer is 0, 2P+1 separated ones otherwise (Okada additivity)."""
    mf = oq.gen_mesh_graded(12e3, 6e3, 1e3, 5, dip, ratio=1.4)
    st = O.gf_ff(mf, LAM, MU, ftype, nrept, beta)
    L = (1.0 + beta) * mf.dx * mf.nx
    a, b = mf.ax[0]
    got, want = st.sum(axis=0), np.zeros((mf.nxi, mf.nxi))
    for j in range(mf.nxi):
        for l in range(mf.nxi):
            spans = [(a + p * L - (mf.nx - 1) * mf.dx, b + p * L) for p in range(-nrept, nrept + 1)]
            if beta == 0.0:
                spans = [(spans[0][0], spans[-1][1])]
            for al in spans:
                u = dc3d(O, mf.x[0], mf.y[j], mf.z[j], dip, ftype, al, mf.axi[l])
                want[j, l] += shear_traction(stress_tensor(u), ftype, dip)
    return relnorm(got, want)


def pin_ff_antiplane(O):
    """Long vertical strike-slip fault, rows graded with depth: summed along
    strike, st is the 2-D strip of screw dislocations [d1, d2] with its
    image, t(ζ) = -μ/(2π) [1/(ζ-d1) - 1/(ζ-d2) - 1/(ζ+d1) + 1/(ζ+d2)]
    (finite-length error ~ (depth/length)^2 ~ 1e-6 here)."""
    mf = oq.gen_mesh_graded(64e3, 8e3, 1e3, 6, 90.0, ratio=1.3)
    S = O.gf_ff(mf, LAM, MU, SS, 10, 0.0).sum(axis=0)
    zeta = -mf.z[:, None]
    d1, d2 = -mf.axi[None, :, 1], -mf.axi[None, :, 0]
    F = 1 / (zeta - d1) - 1 / (zeta - d2) - 1 / (zeta + d1) + 1 / (zeta + d2)
    return relnorm(S, -MU / (2 * np.pi) * F)


def pin_ff_self_stress(O):
    """A slipping patch unloads itself: the K_ff diagonal is negative for
    strike and dip slip at dips 30/60/90 (returns the largest diagonal entry
    over the patch's own scale, which must be < 0)."""
    worst = -np.inf
    for ftype in (SS, DS):
        for dip in (30.0, 60.0, 90.0):
            mf = oq.gen_mesh_graded(12e3, 6e3, 1e3, 5, dip, ratio=1.4)
            st = O.gf_ff(mf, LAM, MU, ftype, 2, 0.0)
            d = np.array([st[0, j, j] for j in range(mf.nxi)])
            worst = max(worst, float(d.max() / np.abs(st).max()))
    return worst


# ---------------- pins of K_mf (GF.jl:123-174) ----------------

MF_CELLS = [(500.0, 3000.0, -6000.0, 800.0, 1200.0, 900.0),
            (-1500.0, -4500.0, -2000.0, 1000.0, 1500.0, 700.0),
            (2500.0, 2500.0, -1200.0, 600.0, 900.0, 500.0)]


def _mf_mesh(dip):
    return oq.gen_mesh_graded(8e3, 4e3, 1e3, 3, dip, ratio=1.5)


def pin_mf_strike_sum(O, ftype, dip, beta, nrept=2):
    """sum over the strike columns of K_mf (Gauss1) = the stress tensor at
    the cell centres from the long rectangle(s) of the images (as for
    K_ff), with every row in the order xx xy xz yy yz zz at (k-1)ne + i and
    column j = (strike q1) + nx (dip q2)."""
    mf, c = _mf_mesh(dip), cells(MF_CELLS)
    K = O.gf_mf(mf, c, LAM, MU, ftype, nrept=nrept, buffer_ratio=beta)
    L = (1.0 + beta) * mf.dx * mf.nx
    a0, b1 = mf.ax[0][0], mf.ax[-1][1]
    got, want = np.zeros((c.ne, mf.nxi, 6)), np.zeros((c.ne, mf.nxi, 6))
    for q2 in range(mf.nxi):
        cols = q2 * mf.nx + np.arange(mf.nx)
        for i in range(c.ne):
            got[i, q2] = [K[k * c.ne + i, cols].sum() for k in range(6)]
            spans = [(a0 + p * L, b1 + p * L) for p in range(-nrept, nrept + 1)]
            if beta == 0.0:
                spans = [(spans[0][0], spans[-1][1])]
            for al in spans:
                u = dc3d(O, c.cx[i], c.cy[i], c.cz[i], dip, ftype, al, mf.axi[q2])
                want[i, q2] += vec6(stress_tensor(u))
    return relnorm(got, want)


def pin_mf_entries(O, ftype, dip):
    """Single entries of K_mf (Gauss1, no images): the stress at the cell
    centre from patch (q1, q2)."""
    mf, c = _mf_mesh(dip), cells(MF_CELLS)
    K = O.gf_mf(mf, c, LAM, MU, ftype, nrept=0)
    got, want = [], []
    for j in (0, 5, mf.nx + 2, 2 * mf.nx + 7):
        q1, q2 = j % mf.nx, j // mf.nx
        for i in range(c.ne):
            u = dc3d(O, c.cx[i], c.cy[i], c.cz[i], dip, ftype, mf.ax[q1], mf.axi[q2])
            want.append(vec6(stress_tensor(u)))
            got.append([K[k * c.ne + i, j] for k in range(6)])
    return relnorm(got, want)


def pin_mf_quadrature(O, ftype, dip):
    """K_mf with the 4^3-point Gauss rule = the cell average of the stress
    (midpoint rule on 6^3 and 12^3 subcells, Richardson-extrapolated)."""
    mf, c = _mf_mesh(dip), cells(MF_CELLS)
    qc, qw = oq.get_quadrature("Gauss6")
    K = O.gf_mf(mf, c, LAM, MU, ftype, qc=qc, qw=qw, nrept=0)
    got, want = [], []
    for j in (0, mf.nx + 3):
        q1, q2 = j % mf.nx, j // mf.nx
        f = lambda x, y, z: vec6(stress_tensor(dc3d(O, x, y, z, dip, ftype, mf.ax[q1], mf.axi[q2])))
        for i in range(c.ne):
            want.append(richardson_average(f, c, i, 6))
            got.append([K[k * c.ne + i, j] for k in range(6)])
    return relnorm(got, want)


# ---------------- Betti reciprocity: K_mf <-> K_fm (GF.jl:123-227) ----------------

def pin_betti_mf_fm(O, ftype, dip):
    """For patch P (unit slip b, area A) and cell C (volume V, unit tensor
    eigenstrain E_p): ∫_C σ_P : E_p dV = ∫_P b·σ_C·n dA.  The left side is
    V W_p K_mf (Gauss6 cell average, no images), the right side the sum of
    K_fm over the sub-patches of P times their areas (midpoint rule on m x m
    sub-patches, m = 4 and 8, Richardson-extrapolated)."""
    mf, c = oq.gen_mesh_graded(4e3, 2e3, 1e3, 2, dip, ratio=1.5), cells(MF_CELLS)
    qc, qw = oq.get_quadrature("Gauss6")
    Kmf = O.gf_mf(mf, c, LAM, MU, ftype, qc=qc, qw=qw, nrept=0)
    V = c.dx * c.dy * c.dz
    lhs = np.array([[V[i] * W_ENG[p] * Kmf[p * c.ne + i, :] for i in range(c.ne)] for p in range(6)])

    def patch_integral(m):
        fine = refine(mf, m)
        Kfm = O.gf_fm(c, fine, LAM, MU, ftype)  # (nf_fine, 6ne)
        area = np.array([(fine.ax[a][1] - fine.ax[a][0]) * (fine.axi[b][1] - fine.axi[b][0])
                         for b in range(fine.nxi) for a in range(fine.nx)])
        out = np.zeros((6, c.ne, mf.nx * mf.nxi))
        for s in range(fine.nx * fine.nxi):
            a, b = s % fine.nx, s // fine.nx
            P = (a // m) + mf.nx * (b // m)
            for p in range(6):
                out[p, :, P] += area[s] * Kfm[s, p * c.ne + np.arange(c.ne)]
        return out

    rhs = (4 * patch_integral(8) - patch_integral(4)) / 3
    return relnorm(lhs, rhs)


def pin_fm_entries(O, ftype, dip):
    """Single entries of K_fm: the shear traction, on the fault plane and in
    the slip direction, of the hex8 stress at the patch centre (x_a, y_b,
    z_b); column (p-1)ne + j (GF.jl:212)."""
    mf, c = oq.gen_mesh_graded(4e3, 2e3, 1e3, 3, dip, ratio=1.5), cells(MF_CELLS)
    K = O.gf_fm(c, mf, LAM, MU, ftype)
    got, want = [], []
    for i in (0, 3, mf.nx + 1, 2 * mf.nx + 2):
        a, b = i % mf.nx, i // mf.nx
        for j in range(c.ne):
            for p in range(6):
                s = hex8_stress(O, mf.x[a], mf.y[b], mf.z[b], c, j, p)
                want.append(shear_traction(tensor6(s), ftype, dip))
                got.append(K[i, p * c.ne + j])
    return relnorm(got, want)


# ---------------- pins of K_mm (GF.jl:250-296) ----------------

MM_CELLS = [(0.0, 0.0, -5000.0, 1000.0, 1200.0, 800.0),
            (2600.0, 500.0, -4200.0, 800.0, 700.0, 1100.0),
            (-900.0, 2800.0, -6500.0, 1300.0, 900.0, 1000.0),
            (1200.0, -2600.0, -2500.0, 700.0, 1000.0, 600.0)]


def pin_mm_reciprocity(O, qtype="Gauss8"):
    """V_j W_k K[(k-1)n + j, (p-1)n + i] = V_i W_p K[(p-1)n + i, (k-1)n + j]
    for distinct, separated cells (cell averages by a 5^3 Gauss rule; the
    residual falls from 1e-4 with 3^3 points to 5e-8 with 5^3)."""
    c = cells(MM_CELLS)
    qc, qw = oq.get_quadrature(qtype)
    K = O.gf_mm_rows(c, c, LAM, MU, qc=qc, qw=qw)
    n = c.ne
    V = c.dx * c.dy * c.dz
    A, B = [], []
    for j in range(n):
        for i in range(n):
            if i == j:
                continue
            for k in range(6):
                for p in range(6):
                    A.append(V[j] * W_ENG[k] * K[k * n + j, p * n + i])
                    B.append(V[i] * W_ENG[p] * K[p * n + i, k * n + j])
    return relnorm(A, B)


def pin_mm_quadrature(O):
    """K_mm rows with the 4^3-point Gauss rule = the receiver-cell average of
    the hex8 stress (midpoint rule on 4^3 and 8^3 subcells, extrapolated)."""
    c = cells(MM_CELLS)
    qc, qw = oq.get_quadrature("Gauss6")
    K = O.gf_mm_rows(c, c, LAM, MU, qc=qc, qw=qw)
    n = c.ne
    got, want = [], []
    for (j, i, p) in ((0, 1, 1), (2, 3, 5), (3, 0, 2)):
        f = lambda x, y, z: hex8_stress(O, x, y, z, c, i, p)
        want.append(richardson_average(f, c, j, 4))
        got.append([K[k * n + j, p * n + i] for k in range(6)])
    return relnorm(got, want)


def slab_stress(p):
    """Inside a laterally infinite layer with uniform eigenstrain E_p in a
    half-space: no lateral strain, σ_xz = σ_yz = σ_zz = 0."""
    E = tensor6(np.eye(6)[p])
    tr = np.trace(E)
    ezz = (LAM * tr + 2 * MU * E[2, 2]) / (LAM + 2 * MU)
    e = np.zeros((3, 3))
    e[2, 2], e[0, 2], e[2, 0], e[1, 2], e[2, 1] = ezz, E[0, 2], E[0, 2], E[1, 2], E[1, 2]
    return vec6(LAM * (np.trace(e) - tr) * np.eye(3) + 2 * MU * (e - E))


def _slab_sums(O, W):
    """Sums over all source cells of the K_mm rows (Gauss1) of the centre
    cell of a W x W layer of 1 km x 1 km x 0.5 km cells at 3 km depth, and of
    a cell 1.5 km above the layer: two (p, k) matrices."""
    m = oq.hex8_box(-W / 2 * 1e3, -W / 2 * 1e3, -3e3, W * 1e3, W * 1e3, -0.5e3, W, W, 1)
    mid = (W // 2) * W + W // 2
    out = []
    for rcv in (SimpleNamespace(**{k: getattr(m, k)[[mid]] for k in _HEX}),
                cells([(0.0, 0.0, -1.5e3, 1e3, 1e3, 0.5e3)])):
        K = O.gf_mm_rows(m, rcv, LAM, MU)
        out.append(np.array([[K[k, p * m.ne:(p + 1) * m.ne].sum() for k in range(6)] for p in range(6)]))
    return out


def pin_mm_slab(O):
    """The layer sums of 61 x 61 and 81 x 81 cells, extrapolated in 1/W (the
    cut-off layer outside W/2 contributes ~ thickness/W at the centre), give
    the slab stress of every unit eigenstrain inside, and nothing 1.5 km
    above the layer."""
    (i61, a61), (i81, a81) = _slab_sums(O, 61), _slab_sums(O, 81)
    inside, above = (81 * i81 - 61 * i61) / 20, (81 * a81 - 61 * a61) / 20
    want = np.array([slab_stress(p) for p in range(6)])
    return max(relnorm(inside, want), float(np.abs(above).max() / (2 * MU)))


def pin_mm_eigvals(O, qtype="Gauss2"):
    """The reference's diagnostic (GF.jl:291-294) on a small graded mantle:
    with cell-averaged rows, diag(V W) K_mm is the interaction energy
    -∫σ:C⁻¹:σ of the eigenstrain fields, symmetric negative semi-definite,
    so K_mm's eigenvalues are real and <= 0.  Returns the larger of max Re
    eig(K_mm) and the largest eigenvalue of the symmetric part of
    diag(V W) K_mm, each over its spectral radius (must be < 0).  (With the
    Gauss1 collocation default the same mesh gives max Re eig = +0.05 of the
    spectral radius: collocation at cell centres is not a cell average.)"""
    m = oq.hex8_box(-4e3, -3e3, -2e3, 8e3, 6e3, -4e3, 4, 3, 2, rfzh=[1.0, 1.6])
    qc, qw = oq.get_quadrature(qtype)
    K = O.gf_mm(m, LAM, MU, qc=qc, qw=qw)
    ev = np.linalg.eigvals(K)
    D = np.concatenate([m.dx * m.dy * m.dz * w for w in W_ENG])
    A = D[:, None] * K
    es = np.linalg.eigvalsh(0.5 * (A + A.T))
    return max(float(ev.real.max() / np.abs(ev).max()), float(es.max() / np.abs(es).max()))


# ---------------- the pins on the oracle ----------------

@pytest.mark.parametrize("ftype,dip", [(SS, 90.0), (SS, 60.0), (DS, 30.0), (DS, 60.0), (DS, 90.0)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_kff_strike_sum(oracle, ftype, dip, beta):
    assert pin_ff_strike_sum(oracle, ftype, dip, beta) <= 1e-10


def test_kff_antiplane_limit(oracle):
    assert pin_ff_antiplane(oracle) <= 1e-5


def test_kff_self_stress_negative(oracle):
    assert pin_ff_self_stress(oracle) < 0.0


def test_unit_slip_jump_sign(oracle):
    """The frame of `frame`: across the plane along n the displacement
    jumps by +b (strike and dip slip, dips 30/60/90)."""
    for dip in (30.0, 60.0, 90.0):
        n, es, ed = frame(dip)
        P = -1.5e3 * ed
        for ftype, b in ((SS, es), (DS, ed)):
            j = dc3d(oracle, *(P + 1e-2 * n), dip, ftype, (-2e3, 2e3), (-3e3, 0.0))[:3] - \
                dc3d(oracle, *(P - 1e-2 * n), dip, ftype, (-2e3, 2e3), (-3e3, 0.0))[:3]
            assert np.abs(j - b).max() < 1e-4


@pytest.mark.parametrize("ftype,dip", [(SS, 90.0), (DS, 60.0)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_kmf_strike_sum(oracle, ftype, dip, beta):
    assert pin_mf_strike_sum(oracle, ftype, dip, beta) <= 1e-10


@pytest.mark.parametrize("ftype,dip", [(SS, 90.0), (DS, 60.0), (DS, 30.0)])
def test_kmf_entries(oracle, ftype, dip):
    assert pin_mf_entries(oracle, ftype, dip) <= 1e-12


@pytest.mark.parametrize("ftype,dip", [(SS, 90.0), (DS, 60.0)])
def test_kmf_quadrature_is_cell_average(oracle, ftype, dip):
    assert pin_mf_quadrature(oracle, ftype, dip) <= 1e-5


@pytest.mark.parametrize("ftype,dip", [(SS, 90.0), (SS, 60.0), (DS, 60.0), (DS, 90.0)])
def test_betti_kmf_kfm(oracle, ftype, dip):
    assert pin_betti_mf_fm(oracle, ftype, dip) <= 1e-4


@pytest.mark.parametrize("ftype,dip", [(SS, 90.0), (DS, 60.0)])
def test_kfm_entries(oracle, ftype, dip):
    assert pin_fm_entries(oracle, ftype, dip) <= 1e-12


def test_kmm_reciprocity(oracle):
    assert pin_mm_reciprocity(oracle) <= 1e-6


def test_kmm_quadrature_is_cell_average(oracle):
    assert pin_mm_quadrature(oracle) <= 1e-5


def test_kmm_slab_sum_rule(oracle):
    assert pin_mm_slab(oracle) <= 4e-3


def test_kmm_eigenvalue_diagnostic(oracle):
    assert pin_mm_eigvals(oracle) < -1e-2


# ---------------- planted mistakes in a rebuilt oracle ----------------

@pytest.fixture(scope="module")
def mutant_factory(oracle, tmp_path_factory):
    """Builds copies of the oracle library with one edit in greens.c; the
    other sources are compiled once."""
    d = tmp_path_factory.mktemp("oracle_mutants")
    cflags = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-D_GNU_SOURCE", "-w"]
    objs = []
    for src in ("okada92.c", "hex8.c", "rhs.c", "tsit5.c", "vcabm.c"):
        o = str(d / (src[:-2] + ".o"))
        subprocess.run(["gcc", *cflags, "-I", ORACLE_DIR, "-c", os.path.join(ORACLE_DIR, src), "-o", o],
                       check=True)
        objs.append(o)
    text = open(os.path.join(ORACLE_DIR, "greens.c")).read()
    count = [0]

    def build(old, new):
        assert text.count(old) >= 1, old
        count[0] += 1
        c = d / ("greens_%d.c" % count[0])
        c.write_text(text.replace(old, new, 1))
        so = str(d / ("liboracle_%d.so" % count[0]))
        subprocess.run(["gcc", *cflags, "-I", ORACLE_DIR, "-shared", "-o", so, str(c), *objs, "-lm"],
                       check=True)
        return so
    return build


@contextmanager
def oracle_lib(O, path):
    """The oracle module with its ctypes library swapped for `path`."""
    saved = O._lib
    O._lib = None
    old_path = O._LIBPATH
    O._LIBPATH = path
    try:
        O.lib()
        yield O
    finally:
        O._lib, O._LIBPATH = saved, old_path


PLANTED = [
    # K_ff
    ("return -sxy * oq_sind(dip) + sxz * oq_cosd(dip);", "return sxy * oq_sind(dip) + sxz * oq_cosd(dip);",
     lambda O: pin_ff_strike_sum(O, SS, 90.0, 0.0) > 1e-10),
    ("(szz - syy) / 2.0 * oq_sind(2.0 * dip)", "(syy - szz) / 2.0 * oq_sind(2.0 * dip)",
     lambda O: pin_ff_strike_sum(O, DS, 60.0, 0.0) > 1e-10),
    ("syz * oq_cosd(2.0 * dip);", "syz * oq_cosd(dip);",
     lambda O: pin_ff_strike_sum(O, DS, 30.0, 0.0) > 1e-10),
    ("lrept = (buffer_ratio + 1.0) * (m->dx * (double)m->nx);",
     "lrept = (buffer_ratio + 1.0) * (m->dx * (double)(m->nx - 1));",
     lambda O: pin_ff_strike_sum(O, SS, 90.0, 0.0) > 1e-10),
    ("for (p = -nrept; p <= nrept; ++p) {\n                    double jump = (double)p * lrept;\n"
     "                    oq_dc3d(alpha, m->x[i]",
     "for (p = -nrept; p < nrept; ++p) {\n                    double jump = (double)p * lrept;\n"
     "                    oq_dc3d(alpha, m->x[i]",
     lambda O: pin_ff_strike_sum(O, SS, 90.0, 0.5) > 1e-10),
    ("st[i + m->nx * j + m->nx * m->nxi * l]", "st[i + m->nx * l + m->nx * m->nxi * j]",
     lambda O: pin_ff_antiplane(O) > 1e-5),
    ("m->ax[0] + jump, m->ax[1] + jump", "m->ax[2] + jump, m->ax[3] + jump",
     lambda O: pin_ff_strike_sum(O, DS, 90.0, 0.5) > 1e-10),
    # K_mf
    ("double rx = ma->cx[i] + qc[3 * w] * ma->dx[i] / 2.0;", "double rx = ma->cx[i] + qc[3 * w] * ma->dx[i];",
     lambda O: pin_mf_quadrature(O, SS, 90.0) > 1e-5),
    ("col[i + ne] += qw[w] * (mu * (u[4] + u[6]));", "col[i + ne] += qw[w] * (mu * (u[4] + u[7]));",
     lambda O: pin_mf_entries(O, SS, 90.0) > 1e-12),
    ("col[i + 4 * ne] += qw[w] * (mu * (u[8] + u[10]));", "col[i + 2 * ne] += qw[w] * (mu * (u[8] + u[10]));",
     lambda O: pin_mf_entries(O, DS, 60.0) > 1e-12),
    ("lrept = (buffer_ratio + 1.0) * (mf->dx * (double)mf->nx);", "lrept = mf->dx * (double)mf->nx;",
     lambda O: pin_mf_strike_sum(O, SS, 90.0, 0.5) > 1e-10),
    ("q1 = j % mf->nx, q2 = j / mf->nx;", "q1 = j / mf->nxi, q2 = j % mf->nxi;",
     lambda O: pin_mf_entries(O, DS, 60.0) > 1e-12),
    # K_fm
    ("if (ftype == 0) return -s[1] * oq_sind(dip) + s[2] * oq_cosd(dip);",
     "if (ftype == 0) return s[1] * oq_sind(dip) + s[2] * oq_cosd(dip);",
     lambda O: pin_betti_mf_fm(O, SS, 90.0) > 1e-4),
    ("return (s[5] - s[3]) / 2.0 * oq_sind(2.0 * dip)", "return (s[3] - s[5]) / 2.0 * oq_sind(2.0 * dip)",
     lambda O: pin_betti_mf_fm(O, DS, 60.0) > 1e-4),
    ("double *col = st + nf * (p * ne + j);", "double *col = st + nf * (j * 6 + p);",
     lambda O: pin_betti_mf_fm(O, SS, 90.0) > 1e-4),
    ("const int64_t a = i % mf->nx, b = i / mf->nx;", "const int64_t a = i / mf->nxi, b = i % mf->nxi;",
     lambda O: pin_fm_entries(O, DS, 60.0) > 1e-12),
    # K_mm
    ("col[k * nr + j] += s[k] * qw[q];", "col[j * 6 + k] += s[k] * qw[q];",
     lambda O: pin_mm_reciprocity(O) > 1e-6),
    ("double rz = rcv->cz[j] + qc[3 * q + 2] * rcv->dz[j] / 2.0;",
     "double rz = rcv->cz[j] + qc[3 * q + 2] * rcv->dz[j];",
     lambda O: pin_mm_quadrature(O) > 1e-5),
    ("double *col = st + r6 * (p * ns + i);", "double *col = st + r6 * (i * 6 + p);",
     lambda O: pin_mm_slab(O) > 4e-3),
]


@pytest.mark.parametrize("k", range(len(PLANTED)))
def test_planted_mistakes_are_caught(oracle, mutant_factory, k):
    old, new, caught = PLANTED[k]
    so = mutant_factory(old, new)
    with oracle_lib(oracle, so) as O:
        assert caught(O), "planted mistake not caught: %r -> %r" % (old, new)
