# FMM4RBHIP.jl -- Julia binding of libfmm4rb_hip.so (gfx950 FMM, C ABI in include/fmm4rb.h).
#
# Adds the algorithm type `FMMHIP` with the reference's (n, N0, eta) fields
# (src/update_particles_field.jl:9-19) plus a precision, and a method of
# `update_particles_field!` for it.  Whenever FMM4RBGPU is loaded -- before or after
# this module, into Main or into any other module -- the method extends
# FMM4RBGPU.update_particles_field! (found by the package's UUID); the
# `Particles`-free entry point `update_fields!` works without it.  `fields` is the
# non-mutating form; whenever ChainRulesCore is loaded it gets an `rrule` whose
# pullback calls the hand-written adjoint (fmm4rb_field_vjp).
#
# Julia is not installed where this library is built and tested; the C ABI is
# what the test suite exercises (tests/test_gpu_parity.py through ctypes).
module FMM4RBHIP

export FMMHIP, update_fields!, fields, field_vjp

const libfmm4rb = get(ENV, "FMM4RB_HIP_LIB",
                      joinpath(@__DIR__, "..", "lib", "libfmm4rb_hip.so"))

const FP64 = Cint(0)
const MIXED = Cint(1)

Base.@kwdef struct FMMHIP{T}
    n::Int
    N0::Int
    eta::T
    precision::Symbol = :fp64   # :fp64 or :mixed
end

mutable struct Context
    ptr::Ptr{Cvoid}
end

function last_error()
    unsafe_string(ccall((:fmm4rb_last_error, libfmm4rb), Cstring, ()))
end

check(rc) = rc == 0 || error("fmm4rb error $rc: $(last_error())")

function Context(device::Integer = 0)
    ref = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fmm4rb_create, libfmm4rb), Cint, (Ptr{Ptr{Cvoid}}, Cint), ref, device))
    ctx = Context(ref[])
    finalizer(c -> ccall((:fmm4rb_destroy, libfmm4rb), Cint, (Ptr{Cvoid},), c.ptr), ctx)
    return ctx
end

const DEFAULT = Ref{Union{Nothing,Context}}(nothing)
default_context() = something(DEFAULT[], (DEFAULT[] = Context(0)))

precision_code(p::Symbol) = p === :fp64 ? FP64 : p === :mixed ? MIXED :
                            throw(ArgumentError("precision must be :fp64 or :mixed"))

"""
    update_fields!(positions, momenta, efields, bfields, charge, alg::FMMHIP; lambda,
                   refit = false)

Vectors of SVector{3,Float64} (isbits, contiguous): the fields are overwritten,
as FMMGPU does (src/update_particles_field.jl:74-75,146-147).  `refit = true`
(tracking steps) keeps the tree and interaction lists of the context's last call
with the same N, N0 and eta and recomputes bounding boxes, p_avg and expansions
(`fmm4rb_field_refit`); rebuild with `refit = false` as particles drift.
"""
function update_fields!(positions::AbstractVector, momenta::AbstractVector,
                        efields::AbstractVector, bfields::AbstractVector,
                        charge::Real, alg::FMMHIP; lambda, ctx::Context = default_context(),
                        refit::Bool = false)
    N = length(positions)
    @assert length(momenta) == N && length(efields) == N && length(bfields) == N
    pos = reinterpret(Float64, positions)
    mom = reinterpret(Float64, momenta)
    E = reinterpret(Float64, efields)
    B = reinterpret(Float64, bfields)
    args = (ctx.ptr, N, pointer(pos), pointer(mom), Float64(charge), Float64(lambda),
            Cint(alg.n), Int64(alg.N0), Float64(alg.eta), precision_code(alg.precision),
            pointer(E), pointer(B), C_NULL)
    GC.@preserve pos mom E B begin
        if refit
            check(ccall((:fmm4rb_field_refit, libfmm4rb), Cint,
                        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint,
                         Int64, Float64, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}), args...))
        else
            check(ccall((:fmm4rb_field, libfmm4rb), Cint,
                        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint,
                         Int64, Float64, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}), args...))
        end
    end
    return nothing
end

"""
    fields(positions, momenta, charge, alg::FMMHIP; lambda) -> (E, B)

Non-mutating form of `update_fields!`: new vectors of the positions' element type.
"""
function fields(positions::AbstractVector, momenta::AbstractVector, charge::Real, alg::FMMHIP;
                lambda, ctx::Context = default_context())
    E = similar(positions)
    B = similar(positions)
    update_fields!(positions, momenta, E, B, charge, alg; lambda = lambda, ctx = ctx)
    return E, B
end

"""
    field_vjp(positions, momenta, Ebar, Bbar, charge, alg::FMMHIP; lambda) -> (gpos, gmom)

Reverse mode of `fields` at fixed tree and interaction lists (rebuilt, not stored):
gpos = sum Ebar.dE/dpos + Bbar.dB/dpos, gmom likewise, in `alg.precision` (the
mixed-precision gradients agree with FP64 to ~2e-7 relative).  `after_forward = true`
(the rrule's pullback) first tries `fmm4rb_field_pullback`, which reuses the tree,
lists, multipoles and local expansions of the context's last forward call when its
inputs and parameters match (about 8% cheaper), and falls back to the self-contained
`fmm4rb_field_vjp`.
"""
function field_vjp(positions::AbstractVector, momenta::AbstractVector, Ebar::AbstractVector,
                   Bbar::AbstractVector, charge::Real, alg::FMMHIP; lambda,
                   ctx::Context = default_context(), after_forward::Bool = false)
    N = length(positions)
    @assert length(momenta) == N && length(Ebar) == N && length(Bbar) == N
    gpos = similar(positions)
    gmom = similar(momenta)
    pos = reinterpret(Float64, positions)
    mom = reinterpret(Float64, momenta)
    eb = reinterpret(Float64, Ebar)
    bb = reinterpret(Float64, Bbar)
    gx = reinterpret(Float64, gpos)
    gp = reinterpret(Float64, gmom)
    args = (ctx.ptr, N, pointer(pos), pointer(mom), Float64(charge), Float64(lambda),
            Cint(alg.n), Int64(alg.N0), Float64(alg.eta), precision_code(alg.precision),
            pointer(eb), pointer(bb), pointer(gx), pointer(gp), C_NULL)
    GC.@preserve pos mom eb bb gx gp begin
        rc = Cint(1)
        if after_forward  # reuse the forward call's local expansions when they still match
            rc = ccall((:fmm4rb_field_pullback, libfmm4rb), Cint,
                       (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint,
                        Int64, Float64, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                        Ptr{Float64}, Ptr{Cvoid}), args...)
        end
        if rc != 0
            check(ccall((:fmm4rb_field_vjp, libfmm4rb), Cint,
                        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Cint,
                         Int64, Float64, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                         Ptr{Float64}, Ptr{Cvoid}), args...))
        end
    end
    return gpos, gmom
end

# The reference package and ChainRulesCore, identified by UUID (FMM4RBGPU's
# Project.toml; ChainRulesCore's registry entry), wherever they are loaded from.
const FMM4RBGPU_ID = Base.PkgId(Base.UUID("6188f75b-cc19-42af-84de-782f8f05992b"), "FMM4RBGPU")
const CHAINRULES_ID = Base.PkgId(Base.UUID("d360d2e6-b24c-11e9-a2a3-2a2ae2dbcce4"),
                                 "ChainRulesCore")
const EXTENDED = Ref(false)
const RRULE_DEFINED = Ref(false)

# update_particles_field!(::FMM4RBGPU.Particles{Float64}, ::FMMHIP; lambda), the drop-in
# for the FMMGPU method (src/update_particles_field.jl:66-148).
function extend_reference!(FM::Module)
    EXTENDED[] && return nothing
    @eval function $(FM).update_particles_field!(particles::$(FM).Particles{Float64},
                                                 alg::FMMHIP; lambda)
        update_fields!(particles.positions, particles.momenta, particles.efields,
                       particles.bfields, particles.charge, alg; lambda = lambda)
    end
    EXTENDED[] = true
    return nothing
end

# ChainRulesCore rule for `fields` (charge, alg and lambda are parameters).
function define_rrule!(CR::Module)
    RRULE_DEFINED[] && return nothing
    @eval function $(CR).rrule(::typeof(fields), positions, momenta, charge, alg::FMMHIP;
                               lambda)
        E, B = fields(positions, momenta, charge, alg; lambda = lambda)
        function fields_pullback(ybar)
            Eb, Bb = $(CR).unthunk(ybar)
            Eb = Eb isa $(CR).AbstractZero ? zero(E) : collect(Eb)
            Bb = Bb isa $(CR).AbstractZero ? zero(B) : collect(Bb)
            gpos, gmom = field_vjp(positions, momenta, Eb, Bb, charge, alg; lambda = lambda,
                                   after_forward = true)
            return ($(CR).NoTangent(), gpos, gmom, $(CR).NoTangent(), $(CR).NoTangent())
        end
        return (E, B), fields_pullback
    end
    RRULE_DEFINED[] = true
    return nothing
end

function on_package_load(id::Base.PkgId)
    id == FMM4RBGPU_ID && extend_reference!(Base.loaded_modules[id])
    id == CHAINRULES_ID && define_rrule!(Base.loaded_modules[id])
    return nothing
end

"""
    update_particles_field!(particles, alg::FMMHIP; lambda)

The reference API through this module's name: forwards to
`FMM4RBGPU.update_particles_field!` (which has the `FMMHIP` method once FMM4RBGPU is
loaded) and fails with an explanation when FMM4RBGPU is not loaded.
"""
function update_particles_field!(particles, alg::FMMHIP; lambda)
    haskey(Base.loaded_modules, FMM4RBGPU_ID) ||
        error("FMM4RBHIP: update_particles_field! needs the FMM4RBGPU package (its " *
              "Particles type); load it with `using FMM4RBGPU`, or call " *
              "update_fields!(positions, momenta, efields, bfields, charge, alg; lambda)")
    FM = Base.loaded_modules[FMM4RBGPU_ID]
    extend_reference!(FM)
    return Base.invokelatest(FM.update_particles_field!, particles, alg; lambda = lambda)
end

function __init__()
    # packages loaded before this module, then every package loaded after it
    for id in (FMM4RBGPU_ID, CHAINRULES_ID)
        haskey(Base.loaded_modules, id) && on_package_load(id)
    end
    push!(Base.package_callbacks, on_package_load)
end

end # module
