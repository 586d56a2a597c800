"""The generated traction pair sums of the hex8 image kernel
(oetqf.jl_amd/csrc/hex8_image_trac0.inc for strike slip, _trac1.inc for dip
slip), and its generated stress combinations (hex8_image_stress.inc: per
moment pair q the diagonal D[k][k] and the symmetric sums D[k][l] + D[l][k]
that a full stress takes, for the K_mm block), against the oracle's
generated 54-term kernel (oracle/hex8_image.inc):
the device quadrature of the fault-traction block K_fm accumulates only
these 12 sums, P[p] and P[6 + p] for eigenstrain pair p, which the generator
simplifies as a whole.  Both bodies are evaluated here statement by
statement at random receiver/source positions; the sums must equal the
oracle's D[0][1] + D[1][0], D[0][2] + D[2][0] (strike slip) and D[2][2] -
D[1][1], D[1][2] + D[2][1] (dip slip) to rounding.  A wrong sign or a
swapped pair moves them by far more than that."""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_INC = os.path.join(ROOT, "oracle", "hex8_image.inc")
PAIRS_INC = {ft: os.path.join(ROOT, "oetqf.jl_amd", "csrc", "hex8_image_trac%d.inc" % ft) for ft in (0, 1)}
STRESS_INC = os.path.join(ROOT, "oetqf.jl_amd", "csrc", "hex8_image_stress.inc")
# per moment pair q: (i, l) entries summed for each of the six combinations
COMBOS = (((0, 0),), ((1, 1),), ((2, 2),), ((0, 1), (1, 0)), ((0, 2), (2, 0)), ((1, 2), (2, 1)))
# (first entry, second entry, sign of the second) of the two sums, as (i, l)
PAIRS = {0: (((0, 1), (1, 0), 1), ((0, 2), (2, 0), 1)),
         1: (((2, 2), (1, 1), -1), ((1, 2), (2, 1), 1))}

_STMT = re.compile(r"^(?:const double (x\d+)|([DP])\[(\d+)\]) = (.*);$")


def _compile(path, text=None, n_out=54, name="D"):
    body = []
    for ln in (open(path).read() if text is None else text).splitlines():
        m = _STMT.match(ln)
        if not m:
            continue
        lhs = m.group(1) or "%s[%s]" % (m.group(2), m.group(3))
        body.append("    %s = %s" % (lhs, m.group(4)))
    src = ("def kernel(X, Y, zr, ys, nu, R):\n    %s = [None] * %d\n" % (name, n_out) + "\n".join(body) +
           "\n    return %s\n" % name)
    env = {"OQ_IPOW": lambda a, n: a ** n}
    exec(compile(src, path, "exec"), env)
    return env["kernel"]


def _sums(D, ft):
    at = lambda il, p: D[il[0] * 18 + il[1] * 6 + p]
    return [at(a, p) + s * at(b, p) for a, b, s in PAIRS[ft] for p in range(6)]


def _combos(D, combos=COMBOS):
    return [sum(D[i * 18 + l * 6 + q] for i, l in c) for q in range(6) for c in combos]


def _agree(f_ref, f_pairs, ft, seed, npts=40, sums=None):
    rng = np.random.default_rng(seed)
    for _ in range(npts):
        X, Y = rng.uniform(-3e4, 3e4, 2)
        zr, ys = -rng.uniform(1e2, 4e4, 2)
        nu = rng.uniform(0.15, 0.35)
        R = math.sqrt(X * X + Y * Y + (zr + ys) ** 2)
        a = np.array((sums or (lambda D: _sums(D, ft)))(f_ref(X, Y, zr, ys, nu, R)))
        b = np.array(f_pairs(X, Y, zr, ys, nu, R))
        scale = np.abs(a).max()
        assert scale > 0
        if np.any(np.abs(a - b) > 1e-10 * np.abs(a) + 1e-13 * scale):
            return False
    return True


@pytest.mark.parametrize("ft", [0, 1])
@pytest.mark.parametrize("seed", [0, 1])
def test_traction_pairs_match_oracle_kernel(ft, seed):
    f_ref = _compile(ORACLE_INC)
    f_pairs = _compile(PAIRS_INC[ft], n_out=12, name="P")
    assert _agree(f_ref, f_pairs, ft, seed)


@pytest.mark.parametrize("ft", [0, 1])
def test_planted_pair_mistakes_are_caught(ft):
    """The oracle side planted wrong: the second sum of the other fault
    type, or the sign of the second entry flipped."""
    f_ref = _compile(ORACLE_INC)
    f_pairs = _compile(PAIRS_INC[ft], n_out=12, name="P")
    good = PAIRS[ft]
    try:
        PAIRS[ft] = (good[0], PAIRS[1 - ft][1])
        assert not _agree(f_ref, f_pairs, ft, 0, npts=10)
        (a, b, s) = good[0]
        PAIRS[ft] = ((a, b, -s), good[1])
        assert not _agree(f_ref, f_pairs, ft, 0, npts=10)
    finally:
        PAIRS[ft] = good


@pytest.mark.parametrize("seed", [0, 1])
def test_stress_combinations_match_oracle_kernel(seed):
    f_ref = _compile(ORACLE_INC)
    f_st = _compile(STRESS_INC, n_out=36, name="P")
    assert _agree(f_ref, f_st, None, seed, sums=_combos)


def test_planted_stress_combination_mistake_is_caught():
    """The oracle side planted wrong: a symmetric sum missing its second entry."""
    f_ref = _compile(ORACLE_INC)
    f_st = _compile(STRESS_INC, n_out=36, name="P")

    def bad(D):
        c = list(COMBOS)
        c[4] = ((0, 2),)
        return _combos(D, tuple(c))
    assert not _agree(f_ref, f_st, None, 0, npts=10, sums=bad)
