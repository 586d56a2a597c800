"""The device copy of the generated hex8 image kernel (oetqf.jl_amd/csrc/
hex8_image.inc) against the oracle's copy (oracle/hex8_image.inc): the device
copy raises every negative power from the reciprocals 1/R and 1/(zr + ys - R)
instead of dividing for each, so both bodies are evaluated here, statement by
statement, at random receiver/source positions and must give the same 54
terms to rounding.  A reciprocal mapped to the wrong base or sign, or a
power off by one, moves terms by far more than that."""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_INC = os.path.join(ROOT, "oracle", "hex8_image.inc")
DEVICE_INC = os.path.join(ROOT, "oetqf.jl_amd", "csrc", "hex8_image.inc")

_STMT = re.compile(r"^(?:const double (x\d+)|D\[(\d+)\]) = (.*);$")


def _compile(path, text=None):
    """The generated body as a Python function of (X, Y, zr, ys, nu, R)."""
    body = []
    for ln in (open(path).read() if text is None else text).splitlines():
        m = _STMT.match(ln)
        if not m:
            continue
        lhs = m.group(1) or "D[%s]" % m.group(2)
        body.append("    %s = %s" % (lhs, m.group(3)))
    src = "def kernel(X, Y, zr, ys, nu, R):\n    D = [None] * 54\n" + "\n".join(body) + "\n    return D\n"
    env = {"OQ_IPOW": lambda a, n: a ** n}
    exec(compile(src, path, "exec"), env)
    return env["kernel"], len(body)


def test_device_copy_has_two_divisions():
    dev = open(DEVICE_INC).read()
    assert "OQ_IPOW(" in dev
    assert not re.search(r"OQ_IPOW\([^)]*, -\d+\)", dev)  # no negative powers left
    assert len(re.findall(r"/", re.sub(r"/\*.*?\*/", "", dev, flags=re.S))) == 2


def _agree(f_ref, f_dev, seed, npts=40):
    """True when every term agrees to 1e-11 of itself (plus 1e-14 of the
    largest term, for terms that cancel to near zero) at every point."""
    rng = np.random.default_rng(seed)
    for _ in range(npts):
        X, Y = rng.uniform(-3e4, 3e4, 2)
        zr, ys = -rng.uniform(1e2, 4e4, 2)
        nu = rng.uniform(0.15, 0.35)
        R = math.sqrt(X * X + Y * Y + (zr + ys) ** 2)
        a = np.array(f_ref(X, Y, zr, ys, nu, R))
        b = np.array(f_dev(X, Y, zr, ys, nu, R))
        scale = np.abs(a).max()
        assert scale > 0
        if np.any(np.abs(a - b) > 1e-11 * np.abs(a) + 1e-14 * scale):
            return False
    return True


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_device_copy_matches_oracle_copy(seed):
    f_ref, n_ref = _compile(ORACLE_INC)
    f_dev, n_dev = _compile(DEVICE_INC)
    assert n_ref == n_dev == 338 + 54
    assert _agree(f_ref, f_dev, seed)


@pytest.mark.parametrize("bad", [
    ("OQ_IPOW((-x18), 3)", "OQ_IPOW(x18, 3)"),    # sign of the negated base's reciprocal
    ("OQ_IPOW(x0, 5)", "OQ_IPOW(x0, 4)"),         # power off by one
    ("OQ_IPOW(x18, 2)", "OQ_IPOW(x0, 2)"),        # reciprocal of the wrong base
    ("*OQ_IPOW(x0, 9)", "*OQ_IPOW(x0, 8)"),       # the rewritten division by R^9
])
def test_planted_rewrite_mistakes_are_caught(bad):
    text = open(DEVICE_INC).read()
    assert bad[0] in text
    f_ref, _ = _compile(ORACLE_INC)
    f_bad, _ = _compile(DEVICE_INC, text.replace(bad[0], bad[1], 1))
    assert not _agree(f_ref, f_bad, 0, npts=10)
