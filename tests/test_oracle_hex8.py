"""Pins of the oracle's hex8 strain-volume stress (oracle/hex8.c).

The reference calls GeoGreensFunctions.stress_vol_hex8! (un-vendored;
src/BEM/GF.jl:215-221, :277-283); no values exist in the reference, so the
restatement is pinned by: the laterally infinite slab limit, a traction-free
free surface, equilibrium, continuity of traction across cuboid faces,
additivity when a cuboid is split, the point-source far field of the
full-space (Kelvin) part, and convergence of the image-part quadrature.
"""
import numpy as np
import pytest

MU, NU = 1.0, 0.25
LAM = 2 * MU * NU / (1 - 2 * NU)
E6 = np.eye(6)


def S(oracle, x, y, z, q, T, L, W, e, nq=6):
    return oracle.stress_vol_hex8(x, y, z, q[0], q[1], q[2], T, L, W, e, MU, NU, nq)


def tensor(s):
    return np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])


@pytest.mark.parametrize("p", range(6))
def test_laterally_infinite_slab(oracle, p):
    """A slab of thickness 1 at depth 2..3, 1e4 wide: the lateral strains
    vanish, so inside sigma = -C:e* with sigma_i3 = 0 (sigma_zz relaxed by
    vertical expansion) and outside sigma = 0 (finite-width error ~1e-4)."""
    q, T, L, W = (0.0, -5e3, -2.0), 1e4, 1e4, 1.0
    e = E6[p]
    et = tensor(e)
    # exact interior: eps = (0,0,ezz) + shear ezx/ezy free, sigma_i3 = 0
    m = LAM * np.trace(et) * np.eye(3) + 2 * MU * et
    want = -m.copy()
    # vertical normal strain relaxes sigma_zz: ezz = m_zz/(lam + 2mu)
    ezz = m[2, 2] / (LAM + 2 * MU)
    want += LAM * ezz * np.eye(3)
    want[2, 2] = 0.0
    want[0, 2] = want[2, 0] = want[1, 2] = want[2, 1] = 0.0
    inside = tensor(S(oracle, 0.3, 0.2, -2.5, q, T, L, W, e))
    above = tensor(S(oracle, 0.3, 0.2, -1.0, q, T, L, W, e))
    below = tensor(S(oracle, 0.3, 0.2, -4.0, q, T, L, W, e))
    assert np.abs(inside - want).max() < 1e-3
    assert np.abs(above).max() < 1e-3 and np.abs(below).max() < 1e-3


@pytest.mark.parametrize("p", range(6))
def test_free_surface_is_traction_free(oracle, rng, p):
    q, T, L, W = (0.0, -0.5, -0.4), 1.0, 1.2, 0.8
    for _ in range(4):
        x, y = rng.uniform(-2, 2, 2)
        s = tensor(S(oracle, x, y, 0.0, q, T, L, W, E6[p]))
        scale = np.abs(tensor(S(oracle, x, y, -0.2, q, T, L, W, E6[p]))).max()
        assert np.abs(s[:, 2]).max() <= 1e-11 * scale


@pytest.mark.parametrize("p", range(6))
def test_equilibrium(oracle, rng, p):
    q, T, L, W = (0.0, -0.5, -1.0), 1.0, 1.0, 1.0
    h = 1e-4
    pts = [(1.7, 0.2, -1.3), (0.1, 0.1, -1.5), (-0.9, 1.4, -0.3)]  # outside, inside, outside
    for (x, y, z) in pts:
        div = np.zeros(3)
        for b, e in enumerate(np.eye(3)):
            sp = tensor(S(oracle, x + h * e[0], y + h * e[1], z + h * e[2], q, T, L, W, E6[p]))
            sm = tensor(S(oracle, x - h * e[0], y - h * e[1], z - h * e[2], q, T, L, W, E6[p]))
            div += (sp[:, b] - sm[:, b]) / (2 * h)
        assert np.abs(div).max() <= 1e-5


@pytest.mark.parametrize("p", range(6))
def test_traction_continuous_across_faces(oracle, p):
    """sigma . n is continuous across each face; the other components jump."""
    q, T, L, W = (0.0, -0.5, -2.0), 1.0, 1.0, 1.0
    eps = 1e-9
    faces = [((0.5, 0.1, -2.4), 0), ((0.2, -0.5, -2.3), 1), ((0.1, 0.2, -3.0), 2)]
    for (x, y, z), ax in faces:
        d = np.zeros(3)
        d[ax] = eps
        a = tensor(S(oracle, x + d[0], y + d[1], z + d[2], q, T, L, W, E6[p]))
        b = tensor(S(oracle, x - d[0], y - d[1], z - d[2], q, T, L, W, E6[p]))
        assert np.abs(a[:, ax] - b[:, ax]).max() <= 1e-6
        on = tensor(S(oracle, x, y, z, q, T, L, W, E6[p]))
        assert np.all(np.isfinite(on)) and np.abs(on[:, ax] - a[:, ax]).max() <= 1e-6


@pytest.mark.parametrize("p", range(6))
def test_additive_when_split(oracle, p):
    q, T, L, W = (0.0, -0.5, -1.5), 1.0, 1.0, 1.2
    for (x, y, z) in [(0.21, 0.1, -2.0), (1.3, -0.7, -0.6), (-0.1, 0.3, -2.9)]:
        whole = S(oracle, x, y, z, q, T, L, W, E6[p])
        # split along x: [qx-T/2, 0.05] and [0.05, qx+T/2]
        x0, x1 = -0.5, 0.05
        a = S(oracle, x, y, z, ((x0 + x1) / 2, q[1], q[2]), x1 - x0, L, W, E6[p])
        b = S(oracle, x, y, z, ((x1 + 0.5) / 2, q[1], q[2]), 0.5 - x1, L, W, E6[p])
        assert np.abs(a + b - whole).max() <= 1e-10 * max(1.0, np.abs(whole).max())


def _kelvin_point_moment_stress(x, m):
    """Full-space stress at x from a point moment m (unit volume):
    u_i = -m_jk dG_ij/dx_k with Kelvin's G, differentiated numerically."""
    def G(r):
        R = np.linalg.norm(r)
        return ((3 - 4 * NU) * np.eye(3) / R + np.outer(r, r) / R**3) / (16 * np.pi * MU * (1 - NU))

    def u(r):
        h = 1e-5
        out = np.zeros(3)
        for k in range(3):
            e = np.zeros(3); e[k] = h
            dG = (G(r + e) - G(r - e)) / (2 * h)
            out -= dG @ m[:, k]
        return out
    h = 1e-4
    gu = np.zeros((3, 3))
    for b in range(3):
        e = np.zeros(3); e[b] = h
        gu[:, b] = (u(x + e) - u(x - e)) / (2 * h)
    eps = 0.5 * (gu + gu.T)
    return LAM * np.trace(eps) * np.eye(3) + 2 * MU * eps


@pytest.mark.parametrize("p", range(6))
def test_far_field_is_point_moment(oracle, p):
    """A small cube deep below the surface seen from afar: the stress tends
    to the Kelvin point-moment field (relative error ~ (size/r)^2)."""
    c = np.array([0.0, 0.0, -1e5])
    a = 0.1
    q = (c[0], c[1] - a / 2, c[2] + a / 2)
    r = np.array([0.6, -0.8, 1.1])
    x = c + r * 5.0
    et = tensor(E6[p])
    m = (LAM * np.trace(et) * np.eye(3) + 2 * MU * et) * a**3
    got = tensor(S(oracle, *x, q, a, a, a, E6[p]))
    want = _kelvin_point_moment_stress(x - c, m)
    assert np.abs(got - want).max() <= 2e-3 * np.abs(want).max()


def test_image_quadrature_converges(oracle, rng):
    q, T, L, W = (0.0, -1.0, -3.0), 2.0, 2.0, 1.5
    for _ in range(4):
        x, y = rng.uniform(-3, 3, 2)
        z = -rng.uniform(0.0, 6.0)
        for p in range(6):
            a = S(oracle, x, y, z, q, T, L, W, E6[p], nq=6)
            b = S(oracle, x, y, z, q, T, L, W, E6[p], nq=9)
            assert np.abs(a - b).max() <= 1e-11 * max(1.0, np.abs(b).max())
