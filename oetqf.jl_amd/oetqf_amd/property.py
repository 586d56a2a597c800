"""Parameter types (src/BEM/property.jl:4-48) with the reference's invariants."""
import numpy as np


class DieterichStateLaw:
    """Aging law dθ/dt = 1 - vθ/L (property.jl:4-5, equation.jl:279)."""


def _arr(x):
    return np.asarray(x, dtype=np.float64)


class RateStateQuasiDynamicProperty:
    """property.jl:10-25; f₀ = 0.6 and v₀ = 1e-6 by default."""

    def __init__(self, a, b, L, σ, η, vpl, f0=0.6, v0=1e-6):
        a, b, L, σ = _arr(a), _arr(b), _arr(L), _arr(σ)
        assert a.shape == b.shape == L.shape == σ.shape, "size(a) == size(b) == size(L) == size(σ)"
        assert f0 > 0, "f₀ > 0"
        assert v0 > 0, "v₀ > 0"
        assert η > 0, "η > 0"
        assert vpl > 0, "vpl > 0"
        self.a, self.b, self.L, self.σ = a, b, L, σ
        self.η, self.vpl, self.f0, self.v0 = float(η), float(vpl), float(f0), float(v0)

    # ASCII aliases
    sigma = property(lambda s: s.σ)
    eta = property(lambda s: s.η)


class DilatancyProperty:
    """property.jl:27-32: per-patch diffusion time tₚ, dilatancy coefficient
    ϵ, gouge compressibility β and ambient pore pressure p₀ (positional order
    as the reference; the Julia field names are aliases)."""

    def __init__(self, tp, eps, beta, p0):
        self.tp, self.eps, self.beta, self.p0 = _arr(tp), _arr(eps), _arr(beta), _arr(p0)
        assert self.tp.shape == self.eps.shape == self.beta.shape == self.p0.shape, \
            "dilatancy property arrays must have the same shape"
        setattr(self, "p\u2080", self.p0)

    # (tₚ is tp after Python's identifier normalization)
    ϵ = property(lambda s: s.eps)
    β = property(lambda s: s.beta)


class PowerLawViscosityProperty:
    """property.jl:34-41: dϵ/dt = γ σ τ^n with n = power - 1."""

    def __init__(self, γ, n, dϵ0):
        γ, n, dϵ0 = _arr(γ).reshape(-1), _arr(n).reshape(-1), _arr(dϵ0).reshape(-1)
        assert len(dϵ0) == 6, "length(dϵ₀) == 6"
        assert len(γ) == len(n), "length(γ) == length(n)"
        self.γ, self.n, self.dϵ0 = γ, n, dϵ0

    gamma = property(lambda s: s.γ)
    deps0 = property(lambda s: s.dϵ0)

    def laws(self):
        return [self]


class CompositePowerLawViscosityProperty:
    """property.jl:43-48: the strain rates of several power laws add up."""

    def __init__(self, piter, dϵ0):
        dϵ0 = _arr(dϵ0).reshape(-1)
        assert len(dϵ0) == 6, "length(dϵ₀) == 6"
        self.piter, self.dϵ0 = list(piter), dϵ0

    deps0 = property(lambda s: s.dϵ0)

    def laws(self):
        return list(self.piter)


def dϵ_dt(p, i, σ, τnorm):
    """Viscosity law (equation.jl:285-292), host evaluation."""
    return sum(law.γ[i] * σ * τnorm ** law.n[i] for law in p.laws())
