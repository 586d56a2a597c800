"""Pins of the oracle's Okada (1992) DC3D (oracle/okada92.c).

The reference holds no dc3d values (GeoGreensFunctions.jl is not vendored;
call sites src/BEM/GF.jl:49-53, :156-160), so the oracle is pinned by what
the mathematics fixes: the gradients are derivatives of the displacement,
the field is in equilibrium, the free surface is traction-free, the
displacement jumps by the unit slip across the fault, the solution is
additive over sub-faults, and a long strike-slip fault reduces to the 2-D
antiplane (screw dislocation) solution.  A wrong sign, term or index in any
of the A/B/C parts fails at least one of these.
"""
import numpy as np
import pytest

ALPHA = 2.0 / 3.0  # lambda = mu
LAM = MU = 1.0
FAULT = dict(depth=4.0, al=(-3.0, 2.0), aw=(-2.0, 1.5))
SLIPS = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
DIPS = [90.0, 0.0, 37.0, 70.0, 135.0]


def U(oracle, x, y, z, dip, d, depth=FAULT["depth"], al=FAULT["al"], aw=FAULT["aw"]):
    u, _ = oracle.dc3d(x, y, z, ALPHA, depth, dip, al[0], al[1], aw[0], aw[1], *d)
    return u


def grad_matrix(u):
    """G[a, b] = d u_a / d x_b from the 12-vector."""
    return np.array([[u[3], u[6], u[9]], [u[4], u[7], u[10]], [u[5], u[8], u[11]]])


def stress(u):
    G = grad_matrix(u)
    e = 0.5 * (G + G.T)
    return LAM * np.trace(e) * np.eye(3) + 2 * MU * e


def fd_grad(oracle, x, y, z, dip, d, h=1e-4):
    g = np.zeros((3, 3))
    for b, e in enumerate(np.eye(3)):
        f = lambda s: U(oracle, x + s * e[0], y + s * e[1], z + s * e[2], dip, d)[:3]
        g[:, b] = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
    return g


@pytest.mark.parametrize("dip", DIPS)
@pytest.mark.parametrize("d", SLIPS)
def test_gradients_are_derivatives_of_displacement(oracle, rng, dip, d):
    worst = 0.0
    for _ in range(12):
        x, y = rng.uniform(-6, 6, 2)
        z = -rng.uniform(0.2, 8.0)
        u = U(oracle, x, y, z, dip, d)
        G = grad_matrix(u)
        worst = max(worst, np.abs(G - fd_grad(oracle, x, y, z, dip, d)).max() / np.abs(G).max())
    assert worst < 1e-7


@pytest.mark.parametrize("dip", DIPS)
@pytest.mark.parametrize("d", SLIPS)
def test_free_surface_is_traction_free(oracle, rng, dip, d):
    for _ in range(10):
        x, y = rng.uniform(-8, 8, 2)
        s0 = stress(U(oracle, x, y, 0.0, dip, d))
        scale = np.abs(stress(U(oracle, x, y, -1.0, dip, d))).max()
        assert np.abs(s0[:, 2]).max() <= 1e-12 * scale


@pytest.mark.parametrize("dip", [90.0, 0.0, 37.0])
@pytest.mark.parametrize("d", SLIPS)
def test_navier_equilibrium(oracle, rng, dip, d):
    h = 1e-4
    for _ in range(6):
        x, y = rng.uniform(-8, 8, 2)
        z = -rng.uniform(0.5, 9.0)
        div = np.zeros(3)
        for b, e in enumerate(np.eye(3)):
            sp = stress(U(oracle, x + h * e[0], y + h * e[1], z + h * e[2], dip, d))
            sm = stress(U(oracle, x - h * e[0], y - h * e[1], z - h * e[2], dip, d))
            div += (sp[:, b] - sm[:, b]) / (2 * h)
        assert np.abs(div).max() <= 1e-5 * np.abs(stress(U(oracle, x, y, z, dip, d))).max()


@pytest.mark.parametrize("dip", [90.0, 0.0, 37.0])
@pytest.mark.parametrize("d", SLIPS)
def test_unit_displacement_jump_across_fault(oracle, dip, d):
    sd, cd = np.sin(np.radians(dip)), np.cos(np.radians(dip))
    xi, eta = -0.5, 0.3
    P = np.array([xi, eta * cd, -FAULT["depth"] + eta * sd])
    n = np.array([0.0, -sd, cd])
    h = 1e-4
    jump = U(oracle, *(P + h * n), dip, d)[:3] - U(oracle, *(P - h * n), dip, d)[:3]
    b = d[0] * np.array([1.0, 0, 0]) + d[1] * np.array([0, cd, sd]) + d[2] * n
    # the jump is +-b (orientation convention) up to O(h)
    assert min(np.abs(jump - b).max(), np.abs(jump + b).max()) < 1e-3


@pytest.mark.parametrize("dip", [90.0, 41.0])
@pytest.mark.parametrize("d", SLIPS)
def test_additive_over_subfaults(oracle, rng, dip, d):
    al, aw = FAULT["al"], FAULT["aw"]
    for _ in range(5):
        x, y = rng.uniform(-6, 6, 2)
        z = -rng.uniform(0.3, 8.0)
        whole = U(oracle, x, y, z, dip, d)
        sx = U(oracle, x, y, z, dip, d, al=(al[0], 0.7)) + U(oracle, x, y, z, dip, d, al=(0.7, al[1]))
        sw = U(oracle, x, y, z, dip, d, aw=(aw[0], -0.4)) + U(oracle, x, y, z, dip, d, aw=(-0.4, aw[1]))
        assert np.abs(sx - whole).max() <= 1e-11 * np.abs(whole).max()
        assert np.abs(sw - whole).max() <= 1e-11 * np.abs(whole).max()


def test_long_strike_slip_fault_is_2d_screw_dislocation(oracle):
    """A vertical strike-slip fault, infinitely long along strike, slipping
    uniformly from the surface to depth W: u_x(y, 0) = b (sign(y)/2 -
    atan(y/W)/pi) (the strip of screw dislocations with its image)."""
    W, L = 2.0, 1e6
    ys = np.array([0.3, 1.0, 2.5, 7.0, -0.8, -4.0])
    got = np.array([U(oracle, 0.0, y, 0.0, 90.0, (1, 0, 0), depth=0.0, al=(-L, L), aw=(-W, 0.0))[0]
                    for y in ys])
    want = np.sign(ys) / 2 - np.arctan(ys / W) / np.pi
    sgn = np.sign(np.dot(got, want))
    assert np.abs(sgn * got - want).max() < 1e-5


def test_receiver_on_fault_edge_returns_zero(oracle):
    """Section 6 of Okada (1992): a receiver on a fault edge is singular and
    the routine returns zeros with a status flag."""
    u, ret = oracle.dc3d(2.0, 0.0, -4.0 + 1.5, ALPHA, 4.0, 90.0, -3.0, 2.0, -2.0, 1.5, 1, 0, 0)
    assert ret == 1 and np.all(u == 0)
    u, ret = oracle.dc3d(0.0, 0.0, 1.0, ALPHA, 4.0, 90.0, -3.0, 2.0, -2.0, 1.5, 1, 0, 0)
    assert ret == 2  # z > 0


def test_sind_cosd_exact_at_right_angles(oracle):
    for a, s, c in [(0, 0, 1), (90, 1, 0), (180, 0, -1), (270, -1, 0), (-90, -1, 0), (450, 1, 0)]:
        assert oracle.sind(a) == s and oracle.cosd(a) == c
    assert oracle.sind(30.0) == pytest.approx(0.5, rel=1e-15)
