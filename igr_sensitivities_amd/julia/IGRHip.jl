# Julia binding of libigr_hip.so (include/igr.h) mirroring the reference's
# IGRAdjoints2D / IGRAdjoints1D entry points.  Untested here: the build
# container has no Julia.  Arrays are the reference's own Float64 arrays
# ((n_p, n_p, N_ex, N_ey) in 2D, (n_p, N_e) in 1D); a state is passed as
# its conserved variables, copied into one contiguous buffer.
module IGRHip

const libigr = joinpath(@__DIR__, "..", "lib", "libigr_hip.so")

const IGR_OK, IGR_EINVAL, IGR_EDOMAIN = 0, 1, 2
const SOLVERS = Dict(:pcg => 0, :jacobi => 1, :chebyshev => 2)
const OBJECTIVES = Dict(:l2_density => 0, :l2_pressure => 1, :kinetic_energy => 2,
                        :windowed_kinetic_energy => 3, :weighted_momentum => 4)

mutable struct Context
    h::Ptr{Cvoid}
    dim::Int
    N::Int
    nv::Int
end

function check(c::Context, rc)
    rc == IGR_OK && return nothing
    msg = unsafe_string(ccall((:igr_last_error, libigr), Cstring, (Ptr{Cvoid},), c.h))
    rc == IGR_EDOMAIN && throw(DomainError(rc, msg))        # Julia's sqrt DomainError
    rc == IGR_EINVAL && error(msg)                           # ErrorException, as the reference
    error("libigr status $rc: $msg")
end

solver_code(s::Symbol) = haskey(SOLVERS, s) ? SOLVERS[s] :
    error("Unknown solver: $s (use :pcg, :jacobi, or :chebyshev)")

function Context(dim, p, nex, ney, lx, ly)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:igr_create, libigr), Cint, (Ptr{Ptr{Cvoid}}, Cint, Cint, Cint, Cint, Cdouble, Cdouble),
               h, dim, p, nex, ney, lx, ly)
    rc == IGR_OK || error("igr_create failed: $rc")
    np = p + 1
    N = dim == 2 ? np * np * nex * ney : np * nex
    c = Context(h[], dim, N, dim == 2 ? 4 : 3)
    finalizer(x -> ccall((:igr_destroy, libigr), Cint, (Ptr{Cvoid},), x.h), c)
    return c
end

Context(basis, mesh) = hasproperty(mesh, :N_ey) ?
    Context(2, basis.p, mesh.N_ex, mesh.N_ey, mesh.Lx, mesh.Ly) : Context(1, basis.p, mesh.N_e, 1, mesh.L, 1.0)

pack(fields...) = reduce(vcat, vec.(fields))
function unpack!(q, fields...)
    N = length(first(fields))
    for (k, f) in enumerate(fields)
        copyto!(f, 1, q, (k - 1) * N + 1, N)
    end
end

set_state!(c, q, Σ=nothing) = check(c, ccall((:igr_set_state, libigr), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                                          c.h, q, Σ === nothing ? C_NULL : Σ))
function get_state!(c, q, Σ)
    check(c, ccall((:igr_get_state, libigr), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), c.h, q, Σ))
end

# run_forward!(μx, μy, ρ, E, Σ, Δt, T, basis, mesh, γ, α, n_iter, solver) (2D/forward.jl:160)
function run_forward!(c::Context, fields_and_Σ, Δt, T, γ, α, n_iter::Int, solver::Symbol=:pcg)
    fields = fields_and_Σ[1:end-1]; Σ = fields_and_Σ[end]
    q = pack(fields...)
    set_state!(c, q)
    ns = Ref{Cint}(0)
    check(c, ccall((:igr_run_forward, libigr), Cint,
                   (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Ptr{Cint}),
                   c.h, Δt, T, γ, α, n_iter, solver_code(solver), ns))
    get_state!(c, q, Σ)
    unpack!(q, fields...)
    return nothing
end

# enzyme_adjoint_ic(μx0, μy0, ρ0, E0, n_steps, Δt, basis, mesh, γ, α; n_iter, solver, objective)
# (2D/enzyme_adjoint.jl:270) -> (J, dμx0, dμy0, dρ0, dE0), plus dJ/dγ, dJ/dα on the GPU path.
function enzyme_adjoint_ic(c::Context, μx0, μy0, ρ0, E0, n_steps, Δt, γ, α;
                           n_iter::Int=10, solver::Symbol=:pcg, objective::Symbol=:l2_density,
                           ckpt_slots::Int=0)
    haskey(OBJECTIVES, objective) || error("Unknown objective: $objective")
    q = pack(μx0, μy0, ρ0, E0)
    set_state!(c, q)
    J = Ref{Cdouble}(0); gb = Ref{Cdouble}(0); ab = Ref{Cdouble}(0)
    qb = similar(q)
    check(c, ccall((:igr_gradient, libigr), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint,
                    Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                   c.h, OBJECTIVES[objective], C_NULL, n_steps, Δt, γ, α, max(n_iter, 1),
                   solver_code(solver), ckpt_slots, J, qb, gb, ab))
    d = [similar(μx0) for _ in 1:4]
    unpack!(qb, d...)
    return J[], d..., gb[], ab[]
end

# ChainRulesCore rrule for the loss J(q0) (ChainRulesCore is not a dependency of
# the reference; load it in the caller's environment to enable this rule):
#
#   function ChainRulesCore.rrule(::typeof(igr_loss), c, μx0, μy0, ρ0, E0, n_steps, Δt, γ, α; kw...)
#       J, dμx, dμy, dρ, dE, dγ, dα = enzyme_adjoint_ic(c, μx0, μy0, ρ0, E0, n_steps, Δt, γ, α; kw...)
#       pullback(J̄) = (NoTangent(), NoTangent(), J̄ .* dμx, J̄ .* dμy, J̄ .* dρ, J̄ .* dE,
#                      NoTangent(), NoTangent(), J̄ * dγ, J̄ * dα)
#       return J, pullback
#   end

end # module
