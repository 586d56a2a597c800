# OetqfAMD.jl -- Julia binding of liboetqf_amd.so (C ABI: include/oetqf_amd.h).
#
# Drop-in methods for Oetqf.jl's hot path on one MI355X:
#   stress_greens_function(...)  -> GPU builders     (src/BEM/GF.jl:31,123,194,250)
#   assemble(...)                -> device RHS handle (src/BEM/equation.jl:81,141)
#   prob.f(du, u, p, t)          -> fused GPU RHS     (src/BEM/equation.jl:156,185)
#   solve(prob, Tsit5(); ...)    -> on-device Tsit5   (OrdinaryDiffEq call sites io.jl:125-130)
# Written against Oetqf v0.3.6 types (RectOkadaMesh, BEMHex8Mesh,
# RateStateQuasiDynamicProperty, PowerLawViscosityProperty).  It needs no
# change to Oetqf itself: `using Oetqf, OetqfAMD` and call OetqfAMD.* .
module OetqfAMD

using Oetqf
using RecursiveArrayTools: ArrayPartition
import HDF5

const libpath = joinpath(@__DIR__, "..", "lib", "liboetqf_amd.so")

struct CFaultMesh
    x::Ptr{Float64}; ax::Ptr{Float64}; y::Ptr{Float64}; z::Ptr{Float64}; axi::Ptr{Float64}
    nx::Int64; nxi::Int64; dx::Float64; dep::Float64; dip::Float64
end
struct CHex8Mesh
    cx::Ptr{Float64}; cy::Ptr{Float64}; cz::Ptr{Float64}
    qx::Ptr{Float64}; qy::Ptr{Float64}; qz::Ptr{Float64}
    dx::Ptr{Float64}; dy::Ptr{Float64}; dz::Ptr{Float64}; ne::Int64
end
struct CQuadrature
    coords::Ptr{Float64}; weights::Ptr{Float64}; nq::Int64
end
struct CRSProps
    a::Ptr{Float64}; b::Ptr{Float64}; L::Ptr{Float64}; sigma::Ptr{Float64}
    eta::Float64; vpl::Float64; f0::Float64; v0::Float64
end
struct CViscProps
    gamma::Ptr{Float64}; n::Ptr{Float64}; nlaws::Int64; deps0::NTuple{6,Float64}
end
struct CSolverOpts
    reltol::Float64; abstol::Float64; dt0::Float64; dtmax::Float64; dtmin::Float64; maxiters::Int64
end

lasterror() = unsafe_string(ccall((:oq_last_error, libpath), Cstring, ()))
function check(rc::Cint)
    rc == 0 && return nothing
    rc == 1 ? throw(ArgumentError(lasterror())) : error("oetqf_amd: " * lasterror())
end

# Vector{NTuple{2,T}} is laid out as 2n contiguous T: pass it as is.
tupleptr(v::Vector{NTuple{2,Float64}}) = Ptr{Float64}(pointer(v))

cmesh(m::Oetqf.RectOkadaMesh) = CFaultMesh(pointer(m.x), tupleptr(m.ax), pointer(m.y), pointer(m.z),
    tupleptr(m.aξ), m.nx, m.nξ, m.Δx, m.dep, m.dip)
cmesh(m::Oetqf.BEMHex8Mesh) = CHex8Mesh(pointer(m.cx), pointer(m.cy), pointer(m.cz),
    pointer(m.qx), pointer(m.qy), pointer(m.qz), pointer(m.Δx), pointer(m.Δy), pointer(m.Δz), length(m.cx))
ftcode(::StrikeSlip) = Cint(0)
ftcode(::DipSlip) = Cint(1)

function quadrature(qtype)
    c, w = Oetqf.get_quadrature(qtype)   # GF.jl:318-328 (gmsh rule or user tuple)
    c = collect(Float64, vec(c)); w = collect(Float64, w)
    return c, w
end

# ---- Green's functions (GF.jl) ----
function stress_greens_function(mesh::Oetqf.RectOkadaMesh, λ::Float64, μ::Float64;
        ftype=StrikeSlip(), fourier::Bool=true, nrept::Integer=2, buffer_ratio::Real=0, kwargs...)
    @assert buffer_ratio ≥ 0 "Argument `buffer_ratio` must be ≥ 0."
    GC.@preserve mesh begin
        cm = Ref(cmesh(mesh))
        if fourier
            out = Array{ComplexF64,3}(undef, mesh.nx, mesh.nξ, mesh.nξ)
            check(ccall((:oq_gf_okada_ff_dft, libpath), Cint,
                (Ref{CFaultMesh}, Float64, Float64, Cint, Int64, Float64, Ptr{ComplexF64}),
                cm, λ, μ, ftcode(ftype), nrept, buffer_ratio, out))
        else
            out = Array{Float64,3}(undef, mesh.nx, mesh.nξ, mesh.nξ)
            check(ccall((:oq_gf_okada_ff, libpath), Cint,
                (Ref{CFaultMesh}, Float64, Float64, Cint, Int64, Float64, Ptr{Float64}),
                cm, λ, μ, ftcode(ftype), nrept, buffer_ratio, out))
        end
    end
    return out
end

function stress_greens_function(mf::Oetqf.RectOkadaMesh, ma::Oetqf.BEMHex8Mesh, λ::Float64, μ::Float64;
        ftype=StrikeSlip(), qtype="Gauss1", nrept::Integer=2, buffer_ratio::Real=0)
    @assert buffer_ratio ≥ 0 "Argument `buffer_ratio` must be ≥ 0."
    c, w = quadrature(qtype)
    out = zeros(Float64, 6length(ma.cx), mf.nx * mf.nξ)
    GC.@preserve mf ma c w begin
        q = Ref(CQuadrature(pointer(c), pointer(w), length(w)))
        check(ccall((:oq_gf_okada_mf, libpath), Cint,
            (Ref{CFaultMesh}, Ref{CHex8Mesh}, Float64, Float64, Cint, Ref{CQuadrature}, Int64, Float64, Ptr{Float64}),
            Ref(cmesh(mf)), Ref(cmesh(ma)), λ, μ, ftcode(ftype), q, nrept, buffer_ratio, out))
    end
    return out
end

function stress_greens_function(ma::Oetqf.BEMHex8Mesh, mf::Oetqf.RectOkadaMesh, λ::Float64, μ::Float64;
        ftype=StrikeSlip())
    out = zeros(Float64, mf.nx * mf.nξ, 6length(ma.cx))
    GC.@preserve mf ma begin
        check(ccall((:oq_gf_hex8_fm, libpath), Cint,
            (Ref{CHex8Mesh}, Ref{CFaultMesh}, Float64, Float64, Cint, Ptr{Float64}),
            Ref(cmesh(ma)), Ref(cmesh(mf)), λ, μ, ftcode(ftype), out))
    end
    return out
end

function stress_greens_function(ma::Oetqf.BEMHex8Mesh, λ::Float64, μ::Float64;
        qtype="Gauss1", checkeigvals::Bool=true)
    c, w = quadrature(qtype)
    out = zeros(Float64, 6length(ma.cx), 6length(ma.cx))
    GC.@preserve ma c w begin
        q = Ref(CQuadrature(pointer(c), pointer(w), length(w)))
        check(ccall((:oq_gf_hex8_mm, libpath), Cint,
            (Ref{CHex8Mesh}, Float64, Float64, Ref{CQuadrature}, Ptr{Float64}),
            Ref(cmesh(ma)), λ, μ, q, out))
    end
    if checkeigvals   # GF.jl:291-294
        maxrλ = maximum(real, Oetqf.eigvals(out))
        println("Maximum real part of eigval is: ", round(maxrλ; digits=4))
    end
    return out
end

# ---- RHS handle (equation.jl) ----
mutable struct GPURHS
    h::Ptr{Cvoid}
    ne::Int
    keep::Any
    function GPURHS(h, ne, keep)
        r = new(h, ne, keep)
        finalizer(x -> (x.h != C_NULL && ccall((:oq_rhs_destroy, libpath), Cint, (Ptr{Cvoid},), x.h); x.h = C_NULL), r)
        return r
    end
end

crs(p::RateStateQuasiDynamicProperty) = CRSProps(pointer(p.a), pointer(p.b), pointer(p.L), pointer(p.σ),
    p.η, p.vpl, p.f₀, p.v₀)
function cvisc(p::PowerLawViscosityProperty, ne)
    γ = collect(Float64, p.γ); n = collect(Float64, p.n)
    return CViscProps(pointer(γ), pointer(n), 1, Tuple(Float64.(p.dϵ₀))), (γ, n)
end
function cvisc(p::CompositePowerLawViscosityProperty, ne)
    γ = reduce(vcat, [collect(Float64, q.γ) for q in p.piter])
    n = reduce(vcat, [collect(Float64, q.n) for q in p.piter])
    return CViscProps(pointer(γ), pointer(n), length(p.piter), Tuple(Float64.(p.dϵ₀))), (γ, n)
end

# Kstorage: Float64, Float32, :bf16, :toeplitz (fault only: K_ff kept as st),
# :fft (K_ff by DFT convolution along strike, equation.jl:44-61, :200) or,
# with assemble_from_mesh only, :fft_grid (:fft, and K_mm as 2D convolutions
# over a mantle whose cells form a uniform horizontal grid)
kcode(T) = T === Float64 ? Cint(0) : T === Float32 ? Cint(1) : T === :toeplitz ? Cint(3) :
           T === :fft ? Cint(5) : T === :fft_grid ? Cint(6) : Cint(2)

struct CDilaProps
    tp::Ptr{Float64}; eps::Ptr{Float64}; beta::Ptr{Float64}; p0::Ptr{Float64}
end

"""assemble(gf, p, u0, tspan; Kstorage=Float64): fault only (equation.jl:81-89)."""
function assemble(gf::AbstractArray, p::RateStateQuasiDynamicProperty, u0::ArrayPartition, tspan::NTuple{2};
        Kstorage=Float64, device::Integer=0, kwargs...)
    nx, nξ = size(u0.x[1])
    isdft = eltype(gf) <: Complex
    g = isdft ? collect(ComplexF64, gf) : collect(Float64, gf)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve p g begin
        check(ccall((:oq_rhs_create, libpath), Cint,
            (Cint, Int64, Int64, Int64, Ref{CRSProps}, Ptr{CViscProps}, Ptr{Cvoid}, Cint,
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Ref{Ptr{Cvoid}}),
            device, nx, nξ, 0, Ref(crs(p)), C_NULL, pointer(g), Cint(isdft), C_NULL, C_NULL, C_NULL,
            kcode(Kstorage), h))
    end
    f = GPURHS(h[], 0, (p,))
    return Oetqf.ODEProblem{true}((du, u, pp, t) -> rhs!(f, du, u, t), u0, tspan, (p, f))
end

"""assemble(gf, p, dila, u0, tspan; Kstorage=Float64): fault only with
dilatancy (equation.jl:108-117), u0 = (v, θ, δ, 𝓅)."""
function assemble(gf::AbstractArray, p::RateStateQuasiDynamicProperty, dila::DilatancyProperty,
        u0::ArrayPartition, tspan::NTuple{2}; Kstorage=Float64, device::Integer=0, kwargs...)
    prob = assemble(gf, p, ArrayPartition(u0.x[1], u0.x[2], u0.x[3]), tspan; Kstorage=Kstorage,
                    device=device)
    f = prob.p[end]
    GC.@preserve dila begin
        d = CDilaProps(pointer(dila.tₚ), pointer(dila.ϵ), pointer(dila.β), pointer(dila.p₀))
        check(ccall((:oq_rhs_set_dilatancy, libpath), Cint, (Ptr{Cvoid}, Ref{CDilaProps}), f.h, Ref(d)))
    end
    f.keep = (f.keep, dila)
    return Oetqf.ODEProblem{true}((du, u, pp, t) -> rhs!(f, du, u, t), u0, tspan, (p, dila, f))
end

"""assemble(gf₁₁, gf₁₂, gf₂₁, gf₂₂, pf, pa, u0, tspan) (equation.jl:141-154)."""
function assemble(gf₁₁::AbstractArray, gf₁₂::AbstractMatrix, gf₂₁::AbstractMatrix, gf₂₂::AbstractMatrix,
        pf::RateStateQuasiDynamicProperty, pa, u0::ArrayPartition, tspan::NTuple{2};
        Kstorage=Float64, device::Integer=0, kwargs...)
    nx, nξ = size(u0.x[1]); ne = size(u0.x[3], 1)
    isdft = eltype(gf₁₁) <: Complex
    g = isdft ? collect(ComplexF64, gf₁₁) : collect(Float64, gf₁₁)
    vs, keepv = cvisc(pa, ne)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve pf keepv g gf₁₂ gf₂₁ gf₂₂ begin
        check(ccall((:oq_rhs_create, libpath), Cint,
            (Cint, Int64, Int64, Int64, Ref{CRSProps}, Ref{CViscProps}, Ptr{Cvoid}, Cint,
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Ref{Ptr{Cvoid}}),
            device, nx, nξ, ne, Ref(crs(pf)), Ref(vs), pointer(g), Cint(isdft),
            pointer(gf₁₂), pointer(gf₂₁), pointer(gf₂₂), kcode(Kstorage), h))
    end
    f = GPURHS(h[], ne, (pf, pa, keepv))
    return Oetqf.ODEProblem{true}((du, u, pp, t) -> rhs!(f, du, u, t), u0, tspan, (pf, pa, f))
end

"""assemble_from_mesh(mf, ma, λ, μ, pf, pa, u0, tspan; ftype, qtype, nrept,
buffer_ratio, Kstorage=Float64): the four stress_greens_function calls
(GF.jl:31-71,123-174,194-227,250-290) and assemble (equation.jl:141-154) in
one step, the blocks built on the device with no host copy of K."""
function assemble_from_mesh(mf::Oetqf.RectOkadaMesh, ma::Oetqf.BEMHex8Mesh, λ::Float64, μ::Float64,
        pf::RateStateQuasiDynamicProperty, pa, u0::ArrayPartition, tspan::NTuple{2};
        ftype=StrikeSlip(), qtype="Gauss1", nrept::Integer=2, buffer_ratio::Real=0, Kstorage=Float64,
        device::Integer=0)
    @assert buffer_ratio ≥ 0 "Argument `buffer_ratio` must be ≥ 0."
    ne = length(ma.cx)
    c, w = quadrature(qtype)
    vs, keepv = cvisc(pa, ne)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve mf ma c w pf keepv begin
        q = Ref(CQuadrature(pointer(c), pointer(w), length(w)))
        check(ccall((:oq_rhs_create_from_mesh, libpath), Cint,
            (Cint, Ref{CFaultMesh}, Ref{CHex8Mesh}, Float64, Float64, Cint, Int64, Float64, Ref{CQuadrature},
             Ref{CRSProps}, Ref{CViscProps}, Cint, Ref{Ptr{Cvoid}}),
            device, Ref(cmesh(mf)), Ref(cmesh(ma)), λ, μ, ftcode(ftype), nrept, buffer_ratio, q,
            Ref(crs(pf)), Ref(vs), kcode(Kstorage), h))
    end
    f = GPURHS(h[], ne, (pf, pa, keepv))
    return Oetqf.ODEProblem{true}((du, u, pp, t) -> rhs!(f, du, u, t), u0, tspan, (pf, pa, f))
end

"""The `ode(du,u,p,t)` contract on host arrays (equation.jl:156-166, :185-205)."""
function rhs!(f::GPURHS, du::ArrayPartition, u::ArrayPartition, t)
    if length(u.x) == 4   # dilatancy (equation.jl:168-183)
        v, θ, δ, 𝓅 = u.x; dv, dθ, dδ, d𝓅 = du.x
        check(ccall((:oq_rhs_eval_dila, libpath), Cint,
            (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            f.h, t, v, θ, δ, 𝓅, dv, dθ, dδ, d𝓅))
    elseif f.ne == 0
        v, θ, δ = u.x; dv, dθ, dδ = du.x
        check(ccall((:oq_rhs_eval, libpath), Cint,
            (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            f.h, t, v, θ, C_NULL, C_NULL, δ, dv, dθ, C_NULL, C_NULL, dδ))
    else
        v, θ, ϵ, σ, δ = u.x; dv, dθ, dϵ, dσ, dδ = du.x
        check(ccall((:oq_rhs_eval, libpath), Cint,
            (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            f.h, t, v, θ, ϵ, σ, δ, dv, dθ, dϵ, dσ, dδ))
    end
    return nothing
end

"""solve on the device; returns (t, u::ArrayPartition, retcode::Symbol).
alg = :Tsit5 (OrdinaryDiffEq's controller) or :VCABM5 (Adams PECE, order <= 5,
the solver of examples/otf-with-mantle.jl:160).  persistent = true runs
whole chunks of steps in one cooperative launch (fault-only, N <= 1024)."""
function solve_gpu(prob; alg=:Tsit5, persistent=false, reltol=1e-6, abstol=1e-8, dt=0.0, dtmax=0.0,
                   dtmin=0.0, maxiters=10^9)
    f = prob.p[end]::GPURHS
    u0 = collect(Float64, prob.u0)                     # packed ArrayPartition order
    s = Ref{Ptr{Cvoid}}(C_NULL)
    alg in (:Tsit5, :VCABM5) || throw(ArgumentError("alg must be :Tsit5 or :VCABM5"))
    if alg == :VCABM5
        check(ccall((:oq_stepper_create_vcabm5, libpath), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Float64, Float64, Ref{CSolverOpts}, Ref{Ptr{Cvoid}}),
            f.h, u0, prob.tspan[1], prob.tspan[2], Ref(CSolverOpts(reltol, abstol, dt, dtmax, dtmin, maxiters)), s))
    else
        check(ccall((:oq_stepper_create, libpath), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Float64, Float64, Ref{CSolverOpts}, Ref{Ptr{Cvoid}}),
            f.h, u0, prob.tspan[1], prob.tspan[2], Ref(CSolverOpts(reltol, abstol, dt, dtmax, dtmin, maxiters)), s))
    end
    if persistent
        st = ccall((:oq_stepper_set_persistent, libpath), Cint, (Ptr{Cvoid}, Cint), s[], 1)
        if st != 0
            msg = lasterror()
            ccall((:oq_stepper_destroy, libpath), Cint, (Ptr{Cvoid},), s[])
            throw(st == 1 ? ArgumentError(msg) : ErrorException(msg))
        end
    end
    rc = Ref{Cint}(0)
    while rc[] == 0
        check(ccall((:oq_stepper_run, libpath), Cint, (Ptr{Cvoid}, Int64, Ref{Cint}), s[], 4096, rc))
    end
    t = Ref{Float64}(0); dtr = Ref{Float64}(0); na = Ref{Int64}(0); nr = Ref{Int64}(0)
    u = similar(u0)
    check(ccall((:oq_stepper_state, libpath), Cint,
        (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ptr{Float64}, Ref{Int64}, Ref{Int64}), s[], t, dtr, u, na, nr))
    ccall((:oq_stepper_destroy, libpath), Cint, (Ptr{Cvoid},), s[])
    out = similar(prob.u0); copyto!(out, u)
    return t[], out, (:Running, :Success, :MaxIters, :DtLessThanMin, :NonFinite)[rc[] + 1]
end

"""wsolve on the device (io.jl:118-133 semantics): every `stride`-th accepted
state is recorded on the GPU, drained every `nstep` snapshots and appended
through `getu(u, t, nothing)` to HDF5 datasets `ustrs` and `tstr` (time is the
trailing, extendible dimension)."""
# The `integrator` argument wsolve_gpu hands to getu at each saved state:
# `it(it.t, Val{1})` is f(u) there (what the reference's handlers read,
# src/io.jl:86, examples/otf-with-mantle.jl:157-159), recorded on the device.
mutable struct GPUSavedState{T}
    t::Float64
    u::T
    du::Union{Nothing,T}
    used::Bool
end
function (s::GPUSavedState)(t, ::Type{Val{1}})
    s.used = true
    t == s.t || throw(ArgumentError("only the derivative at the saved time t = $(s.t) is available"))
    s.du === nothing && throw(ArgumentError("the derivative was not recorded (getu did not ask for it at t0)"))
    return s.du
end

function wsolve_gpu(prob, file, nstep, getu, ustrs, tstr; alg=:Tsit5, stride=1, append=false, force=false,
                    reltol=1e-6, abstol=1e-8, dt=0.0, dtmax=0.0, dtmin=0.0, maxiters=10^9)
    alg in (:Tsit5, :VCABM5) || throw(ArgumentError("alg must be :Tsit5 or :VCABM5"))
    if isfile(file) && !force && !append
        @info "Overwrite existing file $(file) must set `force = true`."
        return nothing
    end
    @assert tstr ∉ ustrs "Duplicate name of $tstr in $ustrs."
    f = prob.p[end]::GPURHS
    u0 = collect(Float64, prob.u0)
    # getu's first call (io.jl:126) also tells whether it reads the derivative
    du0 = similar(prob.u0)
    rhs!(f, du0, prob.u0, prob.tspan[1])
    first = GPUSavedState(prob.tspan[1], prob.u0, du0, false)
    xs0 = getu(prob.u0, prob.tspan[1], first)
    want_du = first.used
    s = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((alg == :VCABM5 ? :oq_stepper_create_vcabm5 : :oq_stepper_create, libpath), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Float64, Float64, Ref{CSolverOpts}, Ref{Ptr{Cvoid}}),
        f.h, u0, prob.tspan[1], prob.tspan[2], Ref(CSolverOpts(reltol, abstol, dt, dtmax, dtmin, maxiters)), s))
    check(ccall((:oq_stepper_record_du, libpath), Cint, (Ptr{Cvoid}, Int64, Int64, Int64, Cint),
        s[], stride, nstep, 0, want_du ? 1 : 0))
    h5 = HDF5.h5open(file, append ? "r+" : "w")
    function save(xs, t)
        @assert length(xs) == length(ustrs) "Unmatched length between solution components and names."
        for (name, x) in zip(ustrs, xs)
            if !haskey(h5, name)
                HDF5.create_dataset(h5, name, Float64, ((size(x)..., 0), (size(x)..., -1)); chunk=(size(x)..., nstep))
            end
            ds = h5[name]; n = size(ds)[end] + 1
            HDF5.set_extent_dims(ds, (size(x)..., n)); ds[ntuple(_ -> Colon(), ndims(x))..., n] = x
        end
        haskey(h5, tstr) || HDF5.create_dataset(h5, tstr, Float64, ((0,), (-1,)); chunk=(nstep,))
        dt_ = h5[tstr]; n = length(dt_) + 1
        HDF5.set_extent_dims(dt_, (n,)); dt_[n] = t
    end
    save(xs0, prob.tspan[1])
    rc = Ref{Cint}(0)
    nsn = Ref{Int64}(0); nst = Ref{Int64}(0)
    snap_t = Vector{Float64}(undef, nstep); snap_u = Vector{Float64}(undef, nstep * length(u0))
    snap_du = want_du ? Vector{Float64}(undef, nstep * length(u0)) : Float64[]
    while rc[] == 0
        check(ccall((:oq_stepper_run, libpath), Cint, (Ptr{Cvoid}, Int64, Ref{Cint}), s[], nstep, rc))
        check(ccall((:oq_stepper_drain_du, libpath), Cint,
            (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Ptr{Float64}, Ref{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            s[], nst, C_NULL, C_NULL, nsn, snap_t, snap_u, want_du ? snap_du : C_NULL))
        for k in 1:nsn[]
            r = (k - 1) * length(u0) + 1:k * length(u0)
            ubuf = similar(prob.u0); copyto!(ubuf, view(snap_u, r))
            dbuf = want_du ? copyto!(similar(prob.u0), view(snap_du, r)) : nothing
            save(getu(ubuf, snap_t[k], GPUSavedState(snap_t[k], ubuf, dbuf, false)), snap_t[k])
        end
        flush(h5)
    end
    close(h5)
    ccall((:oq_stepper_destroy, libpath), Cint, (Ptr{Cvoid},), s[])
    return (:Running, :Success, :MaxIters, :DtLessThanMin, :NonFinite)[rc[] + 1]
end

# ---- ensembles of frictional realizations on one geometry (no reference
# counterpart: the reference solves one remade problem per realization,
# examples/otf-with-mantle.jl:154-161) ----
mutable struct GPUEnsemble
    h::Ptr{Cvoid}
    M::Int
    n::Int
    keep::Any
    function GPUEnsemble(h, M, n, keep)
        r = new(h, M, n, keep)
        finalizer(x -> (x.h != C_NULL && ccall((:oq_ensemble_destroy, libpath), Cint, (Ptr{Cvoid},), x.h);
                        x.h = C_NULL), r)
        return r
    end
end

"""ensemble_gpu(mf, ma, λ, μ, members, pa; Kstorage=:fp16x2): M members
(RateStateQuasiDynamicProperty each, same mesh) sharing the device operator.
Kstorage: Float32, :bf16 or :fp16x2 (FP16 hi/lo split on the FP16 matrix
cores, FP32-class error).  ma = nothing for a fault-only model."""
function ensemble_gpu(mf::Oetqf.RectOkadaMesh, ma, λ::Float64, μ::Float64,
        members::AbstractVector{<:RateStateQuasiDynamicProperty}, pa; Kstorage=:fp16x2,
        ftype=StrikeSlip(), nrept::Integer=2, buffer_ratio::Real=0.0, qtype="Gauss1", device::Integer=0)
    kd = Kstorage === Float32 ? Cint(1) : Kstorage === :bf16 ? Cint(2) : Kstorage === :fp16x2 ? Cint(4) :
         throw(ArgumentError("ensemble Kstorage must be Float32, :bf16 or :fp16x2"))
    rs = [crs(p) for p in members]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    ne = ma === nothing ? 0 : length(ma.cx)
    if ne > 0
        c, w = quadrature(qtype)
        vs, keepv = cvisc(pa, ne)
        GC.@preserve mf ma members c w keepv begin
            check(ccall((:oq_ensemble_create, libpath), Cint,
                (Cint, Ref{CFaultMesh}, Ref{CHex8Mesh}, Float64, Float64, Cint, Int64, Float64,
                 Ref{CQuadrature}, Int64, Ptr{CRSProps}, Ref{CViscProps}, Cint, Ref{Ptr{Cvoid}}),
                device, Ref(cmesh(mf)), Ref(cmesh(ma)), λ, μ, ftcode(ftype), nrept, buffer_ratio,
                Ref(CQuadrature(pointer(c), pointer(w), length(w))), length(members), rs, Ref(vs), kd, h))
        end
        n = 3 * mf.nx * mf.nξ + 12 * ne
    else
        GC.@preserve mf members begin
            check(ccall((:oq_ensemble_create, libpath), Cint,
                (Cint, Ref{CFaultMesh}, Ptr{CHex8Mesh}, Float64, Float64, Cint, Int64, Float64,
                 Ptr{CQuadrature}, Int64, Ptr{CRSProps}, Ptr{CViscProps}, Cint, Ref{Ptr{Cvoid}}),
                device, Ref(cmesh(mf)), C_NULL, λ, μ, ftcode(ftype), nrept, buffer_ratio, C_NULL,
                length(members), rs, C_NULL, kd, h))
        end
        n = 3 * mf.nx * mf.nξ
    end
    return GPUEnsemble(h[], length(members), n, (members, pa))
end

"""ensemble_rhs!(e, dU, U): the RHS of every member; U, dU are n x M
matrices whose columns are packed ArrayPartition states."""
function ensemble_rhs!(e::GPUEnsemble, dU::Matrix{Float64}, U::Matrix{Float64})
    size(U) == (e.n, e.M) && size(dU) == (e.n, e.M) || throw(DimensionMismatch("U, dU must be n x M"))
    check(ccall((:oq_ensemble_eval, libpath), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), e.h, U, dU))
    return dU
end

"""ensemble_solve(e, U0, tspan; reltol, abstol, dt0): lockstep Tsit5 of all
members, each with its own step control.  Returns (t, U, naccept, nreject,
retcodes) per member."""
function ensemble_solve(e::GPUEnsemble, U0::Matrix{Float64}, tspan::NTuple{2}; alg=:Tsit5,
        reltol=1e-6, abstol=1e-8, dt0=1.0, dtmax=0.0, dtmin=0.0, maxiters=10^9, chunk::Integer=4096)
    size(U0) == (e.n, e.M) || throw(DimensionMismatch("U0 must be n x M"))
    alg in (:Tsit5, :VCABM5) || throw(ArgumentError("alg must be :Tsit5 or :VCABM5"))
    opts = CSolverOpts(reltol, abstol, dt0, dtmax, dtmin, maxiters)
    check(ccall((alg == :VCABM5 ? :oq_ensemble_start_vcabm5 : :oq_ensemble_start, libpath), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Float64, Float64, Ref{CSolverOpts}), e.h, U0, tspan[1], tspan[2], Ref(opts)))
    left = Ref{Int64}(e.M)
    while left[] > 0
        check(ccall((:oq_ensemble_run, libpath), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}), e.h, chunk, left))
    end
    t = zeros(e.M); U = similar(U0); na = zeros(Int64, e.M); nr = zeros(Int64, e.M); rc = zeros(Cint, e.M)
    check(ccall((:oq_ensemble_state, libpath), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Int64}, Ptr{Cint}), e.h, t, U, na, nr, rc))
    return t, U, na, nr, [(:Running, :Success, :MaxIters, :DtLessThanMin, :NonFinite)[r + 1] for r in rc]
end

end # module
