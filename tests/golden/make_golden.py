"""Writes tests/golden/*.json: regression vectors of the pinned oracle.

The reference ships no golden values for this path (its tests are
self-consistency checks; GeoGreensFunctions.jl is not vendored), so these
fixtures record the oracle's outputs on fixed inputs after it passed the
mathematical pins of tests/test_oracle_*.py.  They guard both the oracle and
the GPU path against silent drift.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "oetqf.jl_amd"))
import oracle  # noqa: E402
import oetqf_amd as oq  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def dc3d_cases():
    rng = np.random.default_rng(1992)
    out = []
    for dip in (90.0, 70.0, 0.0):
        for slip in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            for _ in range(4):
                x, y = rng.uniform(-6, 6, 2)
                z = -rng.uniform(0.1, 8.0)
                u, ret = oracle.dc3d(x, y, z, 2 / 3, 4.0, dip, -3.0, 2.0, -2.0, 1.5, *slip)
                out.append(dict(x=x, y=y, z=z, alpha=2 / 3, depth=4.0, dip=dip, al=[-3.0, 2.0],
                                aw=[-2.0, 1.5], slip=list(slip), u=u.tolist(), ret=int(ret)))
    return out


def hex8_cases():
    rng = np.random.default_rng(2017)
    out = []
    q, T, L, W = (0.0, -1.0, -3.0), 2.0, 2.0, 1.5
    for _ in range(6):
        x, y = rng.uniform(-3, 3, 2)
        z = -rng.uniform(0.0, 6.0)
        for p in range(6):
            e = np.eye(6)[p]
            s = oracle.stress_vol_hex8(x, y, z, *q, T, L, W, e, 1.0, 0.25, 6)
            out.append(dict(x=x, y=y, z=z, q=list(q), T=T, L=L, W=W, eps=e.tolist(), mu=1.0,
                            nu=0.25, sigma=s.tolist()))
    return out


def kff_case():
    mf = oq.gen_mesh("RectOkada", 20e3, 10e3, 2.5e3, 2.5e3, 41.0)
    st = oracle.gf_ff(mf, 3e10, 3e10, ftype=0, nrept=2, buffer_ratio=1.0)
    return dict(mesh=[20e3, 10e3, 2.5e3, 2.5e3, 41.0], lam=3e10, mu=3e10, ftype=0, nrept=2,
                buffer_ratio=1.0, st_colmajor=st.reshape(-1, order="F").tolist())


if __name__ == "__main__":
    for name, obj in (("dc3d.json", dc3d_cases()), ("hex8.json", hex8_cases()), ("kff.json", kff_case())):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, indent=0)
        print("wrote", name)
