"""Sampled rows of the RHS rebuilt by the oracle from the meshes.

At the workloads' real sizes the oracle cannot afford the whole operator
(12288 x 4096 K_mf alone is 2e8 Okada evaluations with images), so it
assembles only the rows of K that a set of receivers needs -- the Toeplitz
K_ff (GF.jl:31-58), K_fm rows of the sampled patches (GF.jl:194-227), K_mf
and K_mm rows of the sampled cells (GF.jl:123-174, :250-290) -- and
evaluates those rows of the coupled or fault-only `ode` (equation.jl:156-166,
:185-205) for any member's properties and state.
"""
from types import SimpleNamespace

import numpy as np

_HEX = ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")


class OracleRows:
    def __init__(self, oracle, mf, ma, lam, mu, I, J, C=(), nrept=2, buffer_ratio=0.0):
        self.o = oracle
        self.mf, self.ma = mf, ma
        self.I, self.J = np.asarray(I), np.asarray(J)
        self.C = np.asarray(C, dtype=np.int64)
        self.st = oracle.gf_ff(mf, lam, mu, nrept=nrept, buffer_ratio=buffer_ratio)
        self.kfm = self.kmf = self.kmm = None
        if ma is not None:
            sub = SimpleNamespace(x=mf.x[self.I], ax=mf.ax[self.I], y=mf.y[self.J],
                                  z=mf.z[self.J], axi=mf.axi[self.J], nx=len(self.I),
                                  nxi=len(self.J), dx=mf.dx, dep=mf.dep, dip=mf.dip)
            self.kfm = oracle.gf_fm(ma, sub, lam, mu)  # row a + len(I) b
            if len(self.C):
                msub = SimpleNamespace(**{k: getattr(ma, k)[self.C] for k in _HEX})
                self.kmf = oracle.gf_mf(mf, msub, lam, mu, nrept=nrept,
                                        buffer_ratio=buffer_ratio)  # rows k len(C) + c
                self.kmm = oracle.gf_mm_rows(ma, msub, lam, mu)

    def rows(self, pf, pa, v, theta, sigma=None):
        """(dv at patches I x J (Fortran-flattened), dσ at cells C as
        (6, len(C)) component-major) for one member."""
        o, mf = self.o, self.mf
        s_f = v - pf.vpl
        s_m = None
        if self.ma is not None:
            de = o.update_strain_rate(pa, sigma)
            s_m = (de - np.asarray(pa.deps0)[None, :]).reshape(-1, order="F")
        nI, nJ = len(self.I), len(self.J)
        dtau = np.zeros((nI, nJ))
        for a, i in enumerate(self.I):
            D = np.abs(i - np.arange(mf.nx))
            for b, j in enumerate(self.J):
                dtau[a, b] = np.sum(self.st[D, j, :] * s_f)
                if s_m is not None:
                    dtau[a, b] += self.kfm[a + nI * b] @ s_m
        ix = np.ix_(self.I, self.J)
        psub = SimpleNamespace(a=pf.a[ix], b=pf.b[ix], L=pf.L[ix], sigma=pf.sigma[ix],
                               eta=pf.eta, vpl=pf.vpl, f0=pf.f0, v0=pf.v0)
        dv, _, _ = o.update_fault(psub, v[ix], theta[ix], dtau)
        ds = None
        if self.kmf is not None:
            ds = (self.kmf @ s_f.reshape(-1, order="F") + self.kmm @ s_m).reshape(6, len(self.C))
        return dv, ds

    def pick(self, dv_full, ds_full=None):
        """The same rows out of a device result (dv (nx, nξ), dσ (ne, 6))."""
        dv = dv_full[np.ix_(self.I, self.J)].reshape(-1, order="F")
        ds = None if ds_full is None or not len(self.C) else ds_full[self.C, :].T
        return dv, ds
