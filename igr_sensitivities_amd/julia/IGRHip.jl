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
                        :windowed_kinetic_energy => 3, :weighted_momentum => 4,
                        :windowed_density => 5, :disc_kinetic_energy => 6)
const OPT_GRAPHS, OPT_TILE, OPT_HTILE, OPT_AUTOTUNE, OPT_RTILE, OPT_GROUPS, OPT_WAVE1D, OPT_KEEP, OPT_RGROUPS = 0, 1, 2, 3, 4, 5, 6, 7, 8

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

# The reference's IC builders (IC_BUILDERS, build_oseen_pair_blast_ic!, the
# init_* functions; 2D/initial_conditions.jl, 1D/initial_conditions.jl)
# evaluated on the GPU.  build_ic!(c, (μx0, μy0, ρ0, E0), :taylor_green, A, γ)
# fills the caller's arrays as builder!(μx0, μy0, ρ0, E0, A, mesh, γ) does
# and leaves the state in the context; prm overrides the keyword defaults in
# the order of include/igr.h.  dA: optional arrays receiving ∂q0/∂A.
const IC_TYPES = Dict(:constant => 0, :isentropic_vortex => 1, :taylor_green => 2, :sedov_blast => 3,
                      :oseen_pair_blast => 4, :init_entropy_wave => 5, :acoustic_pulse => 6, :entropy_wave => 7,
                      :density_pulse => 8, :pressure_pulse => 9, :velocity_pulse => 10)
function build_ic!(c::Context, fields::Tuple, ic_type::Symbol, A, γ; prm=Float64[], dA=nothing)
    haskey(IC_TYPES, ic_type) || error("Unknown IC type: $ic_type")
    p = Vector{Float64}(prm)
    d = dA === nothing ? C_NULL : Vector{Float64}(undef, c.nv * c.N)
    check(c, ccall((:igr_build_ic, libigr), Cint,
                   (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Ptr{Cdouble}, Cint, Ptr{Cdouble}),
                   c.h, IC_TYPES[ic_type], A, γ, isempty(p) ? C_NULL : p, length(p), d))
    q = Vector{Float64}(undef, c.nv * c.N)
    get_state!(c, q, C_NULL)
    unpack!(q, fields...)
    dA === nothing || unpack!(d, dA...)
    return nothing
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

# run_forward_store(q0..., Δt, T, basis, mesh, γ, α; n_iter, solver)
# (2D/forward.jl:111; 1D/forward.jl:137) -> (snapshots, n_steps).  The records
# stay in HBM; `Snapshots` indexes them like the reference's vector of tuples.
struct Snapshots
    c::Context
    n_steps::Int
end
Base.length(s::Snapshots) = s.n_steps + 1
function Base.getindex(s::Snapshots, k::Int)        # snapshots[k], k = 1 .. n_steps + 1
    1 <= k <= length(s) || throw(BoundsError(s, k))
    q = Vector{Float64}(undef, s.c.nv * s.c.N); Σ = Vector{Float64}(undef, s.c.N)
    check(s.c, ccall((:igr_get_snapshot, libigr), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}),
                     s.c.h, k - 1, q, Σ))
    return (ntuple(v -> q[(v - 1) * s.c.N + 1:v * s.c.N], s.c.nv)..., Σ)   # flat fields; reshape as needed
end

run_forward_store(c::Context, μx0, μy0, ρ0, E0, Δt, T, γ, α; n_iter::Int=10, solver::Symbol=:pcg) =
    _run_forward_store(c, pack(μx0, μy0, ρ0, E0), Δt, T, γ, α, n_iter, solver)      # 2D
run_forward_store(c::Context, μ0, ρ0, E0, Δt, T, γ, α; n_iter::Int=10, solver::Symbol=:pcg) =
    _run_forward_store(c, pack(μ0, ρ0, E0), Δt, T, γ, α, n_iter, solver)            # 1D
function _run_forward_store(c::Context, q0, Δt, T, γ, α, n_iter, solver)
    set_state!(c, q0)
    ns = Ref{Cint}(0)
    check(c, ccall((:igr_run_forward_store, libigr), Cint,
                   (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Ptr{Cint}),
                   c.h, Δt, T, γ, α, n_iter, solver_code(solver), ns))
    return Snapshots(c, ns[]), Int(ns[])
end

# run_adjoint_conservative(snapshots, n_steps, Δt, T, basis, mesh, γ, α; n_iter, objective, solver)
# (2D/adjoint_pde.jl:522; 1D/adjoint_pde.jl:466) -> the adjoint fields at t = 0.
function run_adjoint_conservative(s::Snapshots, n_steps, Δt, T, γ, α;
                                  n_iter::Int=10, objective::Symbol=:l2_density, solver::Symbol=:pcg)
    haskey(OBJECTIVES, objective) || error("Unknown objective: $objective")
    c = s.c
    a = Vector{Float64}(undef, c.nv * c.N)
    check(c, ccall((:igr_run_adjoint_conservative, libigr), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint,
                    Ptr{Cdouble}),
                   c.h, OBJECTIVES[objective], C_NULL, n_steps, Δt, T, γ, α, n_iter, solver_code(solver), a))
    return ntuple(v -> a[(v - 1) * c.N + 1:v * c.N], c.nv)
end

# 1D enzyme_adjoint_ic (1D/enzyme_adjoint.jl:255-314): the gradient of the
# fixed-step objective, mapped to primitive adjoints as the reference does
# (:306-311).  Returns (J, adj_ρ, adj_u, adj_p), plus dJ/dγ, dJ/dα.
function enzyme_adjoint_ic_1d(c::Context, μ0, ρ0, E0, n_steps, Δt, γ, α;
                              n_iter::Int=10, solver::Symbol=:pcg, objective::Symbol=:l2_density,
                              ckpt_slots::Int=0)
    haskey(OBJECTIVES, objective) || error("Unknown objective: $objective (use :l2_density or :l2_pressure)")
    q = pack(μ0, ρ0, E0)
    set_state!(c, q)
    J = Ref{Cdouble}(0); gb = Ref{Cdouble}(0); ab = Ref{Cdouble}(0)
    qb = similar(q)
    check(c, ccall((:igr_gradient, libigr), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint,
                    Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                   c.h, OBJECTIVES[objective], C_NULL, n_steps, Δt, γ, α, max(n_iter, 1),
                   solver_code(solver), ckpt_slots, J, qb, gb, ab))
    dμ, dρ, dE = similar(μ0), similar(ρ0), similar(E0)
    unpack!(qb, dμ, dρ, dE)
    u0 = μ0 ./ ρ0
    adj_ρ = dρ .+ u0 .* dμ .+ (u0 .^ 2 ./ 2) .* dE
    adj_u = ρ0 .* dμ .+ ρ0 .* u0 .* dE
    adj_p = dE ./ (γ - 1)
    return J[], adj_ρ, adj_u, adj_p, gb[], ab[]
end

# run_sensitivity(q0..., s0..., Δt, T, basis, mesh, γ, α; n_iter, solver[, elliptic_rhs, sγ])
# (2D/sensitivity.jl:353; 1D/sensitivity.jl:475): q0 and s0 are the packed
# conserved fields; returns (q, Σ, s, ς) at T.
function run_sensitivity(c::Context, q0::Vector{Float64}, s0::Vector{Float64}, Δt, T, γ, α;
                         n_iter::Int=10, solver::Symbol=:pcg, elliptic_rhs::Symbol=:form2, sγ=0.0)
    form = elliptic_rhs == :form1 ? 1 : elliptic_rhs == :form2 ? 2 :
           error("Unknown elliptic_rhs: $elliptic_rhs (use :form1 or :form2)")
    set_state!(c, q0)
    s = similar(s0); ς = zeros(c.N); ns = Ref{Cint}(0)
    check(c, ccall((:igr_run_sensitivity, libigr), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint, Cdouble,
                    Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cint}),
                   c.h, 1, s0, Δt, T, γ, α, n_iter, solver_code(solver), form, sγ, s, ς, ns))
    q = similar(q0); Σ = zeros(c.N)
    get_state!(c, q, Σ)
    return q, Σ, s, ς
end

# Batched ensembles: one igr_batch_run_sensitivity call for many contexts
# (struct igr_batch_item of include/igr.h, field for field).
struct BatchItem
    ctx::Ptr{Cvoid}
    dt::Cdouble; t_end::Cdouble; gamma::Cdouble; alpha::Cdouble; sgam::Cdouble
    n_tan::Cint
    s0::Ptr{Cdouble}; s_out::Ptr{Cdouble}; vs_out::Ptr{Cdouble}
    n_steps::Cint; status::Cint
end

# runs: vector of (ctx, q0, s0 (n_tan*nv*N, may be empty), Δt, T, γ, α, sγ)
function run_sensitivity_batch(runs; n_iter::Int=10, solver::Symbol=:pcg, elliptic_rhs::Symbol=:form2)
    form = elliptic_rhs == :form1 ? 1 : 2
    outs = [(similar(r[3]), zeros(length(r[3]) == 0 ? 0 : div(length(r[3]), r[1].nv))) for r in runs]
    for r in runs
        set_state!(r[1], r[2])
    end
    items = [BatchItem(r[1].h, r[4], r[5], r[6], r[7], r[8], div(length(r[3]), r[1].nv * r[1].N),
                       length(r[3]) == 0 ? C_NULL : pointer(r[3]),
                       length(r[3]) == 0 ? C_NULL : pointer(o[1]), length(r[3]) == 0 ? C_NULL : pointer(o[2]),
                       0, 0) for (r, o) in zip(runs, outs)]
    rc = GC.@preserve runs outs ccall((:igr_batch_run_sensitivity, libigr), Cint,
                                      (Ptr{BatchItem}, Cint, Cint, Cint, Cint),
                                      items, length(items), n_iter, solver_code(solver), form)
    rc == IGR_OK || check(runs[1][1], rc)
    return outs
end

# Tile shapes (2D, p = 3): set_option!(c, OPT_TILE, 804), or autotune!(c) on the current state.
set_option!(c::Context, opt, value) =
    check(c, ccall((:igr_set_option, libigr), Cint, (Ptr{Cvoid}, Cint, Cint), c.h, opt, value))
autotune!(c::Context) = set_option!(c, OPT_AUTOTUNE, 0)

# ---------------------------------------------------------------------------
# The reference's own signatures (basis and mesh arguments, no context), so a
# paper script's hot-path calls run unchanged on the GPU path: take these
# names from IGRHip.D2 / IGRHip.D1 instead of IGRAdjoints2D / IGRAdjoints1D
# (types, meshes and plotting still come from the reference).  One context
# per mesh is created on first use and kept; results come back shaped like
# the inputs.  The ChainRulesCore rule for D2.loss is in IGRHipChainRules.jl.
const _CTX = Dict{Any, Context}()
ctx_for(basis, mesh) = get!(() -> Context(basis, mesh), _CTX, objectid(mesh))
ctx_key(key, make) = get!(make, _CTX, key)
shaped(c::Context, q, like) = ntuple(v -> reshape(q[(v - 1) * c.N + 1:v * c.N], size(like)), c.nv)

function run_steps_objective(c::Context, q0, n_steps, Δt, γ, α, n_iter, solver, objective)
    haskey(OBJECTIVES, objective) || error("Unknown objective: $objective")
    set_state!(c, q0)
    check(c, ccall((:igr_run_steps, libigr), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cint, Cint),
                   c.h, n_steps, Δt, γ, α, max(n_iter, 1), solver_code(solver)))
    J = Ref{Cdouble}(0)
    check(c, ccall((:igr_objective, libigr), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cdouble, Ptr{Cdouble}),
                   c.h, OBJECTIVES[objective], C_NULL, γ, J))
    return J[]
end

module D2   # IGRAdjoints2D's hot-path names
import ..IGRHip
import ..IGRHip: Context, Snapshots, ctx_for, ctx_key, shaped, pack, unpack!, set_state!, get_state!
export run_forward!, run_forward, run_forward_store, run_adjoint_conservative, enzyme_adjoint_ic, run_sensitivity,
       loss, init_constant_state, init_isentropic_vortex, build_taylor_green_ic!, build_sedov_blast_ic!,
       build_oseen_pair_blast_ic!, IC_BUILDERS

ctx_mesh(mesh) = ctx_key(objectid(mesh), () -> Context(2, size(mesh.x, 1) - 1, mesh.N_ex, mesh.N_ey, mesh.Lx, mesh.Ly))

# forward.jl:41
run_forward!(μx, μy, ρ, E, Σ, Δt, T, basis, mesh, γ, α, n_iter::Int, solver::Symbol=:pcg) =
    IGRHip.run_forward!(ctx_for(basis, mesh), (μx, μy, ρ, E, Σ), Δt, T, γ, α, n_iter, solver)
# forward.jl:84, 91
function run_forward(μx0, μy0, ρ0, E0, Δt, T, basis, mesh, γ, α, n_iter::Int, solver::Symbol=:pcg)
    μx, μy, ρ, E = copy(μx0), copy(μy0), copy(ρ0), copy(E0)
    Σ = zeros(size(μx))
    run_forward!(μx, μy, ρ, E, Σ, Δt, T, basis, mesh, γ, α, n_iter, solver)
    return μx, μy, ρ, E, Σ
end
run_forward(μx0, μy0, ρ0, E0, Δt, T, basis, mesh, γ, α; n_iter=10, solver=:pcg) =
    run_forward(μx0, μy0, ρ0, E0, Δt, T, basis, mesh, γ, α, Int(n_iter), solver)
# forward.jl:111 -> (snapshots, n_steps); the records stay in HBM
run_forward_store(μx0, μy0, ρ0, E0, Δt, T, basis, mesh, γ, α; n_iter=10, solver=:pcg) =
    IGRHip.run_forward_store(ctx_for(basis, mesh), μx0, μy0, ρ0, E0, Δt, T, γ, α; n_iter=Int(n_iter), solver=solver)
# adjoint_pde.jl:522 -> (aμx, aμy, aρ, aE)
function run_adjoint_conservative(snapshots::Snapshots, n_steps, Δt, T, basis, mesh, γ, α;
                                  n_iter=10, objective=:l2_density, solver=:pcg)
    a = IGRHip.run_adjoint_conservative(snapshots, n_steps, Δt, T, γ, α; n_iter=Int(n_iter), objective=objective,
                                        solver=solver)
    return ntuple(v -> reshape(a[v], size(mesh.x)), 4)
end
# enzyme_adjoint.jl:270 -> (J, dμx0, dμy0, dρ0, dE0)
function enzyme_adjoint_ic(μx0, μy0, ρ0, E0, n_steps, Δt, basis, mesh, γ, α;
                           n_iter::Int=10, solver::Symbol=:pcg, objective::Symbol=:l2_density)
    r = IGRHip.enzyme_adjoint_ic(ctx_for(basis, mesh), μx0, μy0, ρ0, E0, n_steps, Δt, γ, α;
                                 n_iter=n_iter, solver=solver, objective=objective)
    return r[1:5]
end
# J of the Enzyme objectives (enzyme_adjoint.jl:50-76): the primal of the ChainRulesCore rule
loss(μx0, μy0, ρ0, E0, n_steps, Δt, basis, mesh, γ, α; n_iter::Int=10, solver::Symbol=:pcg,
     objective::Symbol=:l2_density) =
    IGRHip.run_steps_objective(ctx_for(basis, mesh), pack(μx0, μy0, ρ0, E0), n_steps, Δt, γ, α, n_iter, solver,
                               objective)
# sensitivity.jl:353 -> (μx, μy, ρ, E, Σ, ζx, ζy, σ, η, ς)
function run_sensitivity(μx0, μy0, ρ0, E0, ζx0, ζy0, σ0, η0, Δt, T, basis, mesh, γ, α; n_iter=10, solver=:pcg)
    c = ctx_for(basis, mesh)
    q, Σ, s, ς = IGRHip.run_sensitivity(c, pack(μx0, μy0, ρ0, E0), pack(ζx0, ζy0, σ0, η0), Δt, T, γ, α;
                                        n_iter=Int(n_iter), solver=solver)
    sz = size(μx0)
    return (shaped(c, q, μx0)..., reshape(Σ, sz), shaped(c, s, μx0)..., reshape(ς, sz))
end

# initial_conditions.jl:18-232, built on the GPU
function init_constant_state(mesh, γ; ρ0=1.0, ux0=1.0, uy0=1.0, p0=1.0)
    f = ntuple(_ -> zeros(size(mesh.x)), 4)
    IGRHip.build_ic!(ctx_mesh(mesh), f, :constant, 0.0, γ; prm=[ρ0, ux0, uy0, p0])
    return f
end
function init_isentropic_vortex(mesh, γ; β=5.0, ux0=1.0, uy0=1.0)
    f = ntuple(_ -> zeros(size(mesh.x)), 4)
    IGRHip.build_ic!(ctx_mesh(mesh), f, :isentropic_vortex, 0.0, γ; prm=[β, ux0, uy0])
    return f
end
build_taylor_green_ic!(μx0, μy0, ρ0, E0, A, mesh, γ) =
    IGRHip.build_ic!(ctx_mesh(mesh), (μx0, μy0, ρ0, E0), :taylor_green, A, γ)
build_sedov_blast_ic!(μx0, μy0, ρ0, E0, A, mesh, γ; p_bg=1.0, σ=0.1) =
    IGRHip.build_ic!(ctx_mesh(mesh), (μx0, μy0, ρ0, E0), :sedov_blast, A, γ; prm=[p_bg, σ])
function build_oseen_pair_blast_ic!(μx0, μy0, ρ0, E0, A, mesh, γ; a_core=0.05, M_star=0.5,
                                    vort_centers=((0.45, 0.60, +1), (0.45, 0.40, -1)),
                                    blast_xc=0.25, blast_yc=0.50, σ=0.09)
    prm = Float64[a_core, M_star, blast_xc, blast_yc, σ]
    for (xc, yc, sgn) in vort_centers
        append!(prm, (xc, yc, sgn))
    end
    IGRHip.build_ic!(ctx_mesh(mesh), (μx0, μy0, ρ0, E0), :oseen_pair_blast, A, γ; prm=prm)
end
const IC_BUILDERS = Dict{Symbol, Tuple{Function, String}}(
    :taylor_green => (build_taylor_green_ic!, "ux=A sin cos, uy=-A cos sin, ρ=1, p=1"),
    :sedov_blast  => (build_sedov_blast_ic!,  "ρ=1, u=0, p=p_bg+A exp(-r²/σ²)"),
)
end # module D2

module D1   # IGRAdjoints1D's hot-path names
import ..IGRHip
import ..IGRHip: Context, Snapshots, ctx_for, ctx_key, shaped, pack, unpack!, set_state!, get_state!
export run_forward!, run_forward, run_forward_store, run_adjoint_conservative, enzyme_adjoint_ic, run_sensitivity,
       init_constant_state, init_entropy_wave, build_acoustic_pulse_ic!, build_entropy_wave_ic!,
       build_density_pulse_ic!, build_pressure_pulse_ic!, build_velocity_pulse_ic!, IC_BUILDERS

ctx_x(mesh_x, L) = ctx_key((size(mesh_x), L), () -> Context(1, size(mesh_x, 1) - 1, size(mesh_x, 2), 1, L, 1.0))

# forward.jl:50
run_forward!(μ, ρ, E, Σ, Δt, T, basis, mesh, γ, α, n_iter::Int, solver::Symbol=:pcg) =
    IGRHip.run_forward!(ctx_for(basis, mesh), (μ, ρ, E, Σ), Δt, T, γ, α, n_iter, solver)
# forward.jl:97, 110
function run_forward(μ0, ρ0, E0, Δt, T, basis, mesh, γ, α, n_iter::Int, solver::Symbol=:pcg)
    μ, ρ, E = copy(μ0), copy(ρ0), copy(E0)
    Σ = zeros(size(μ))
    run_forward!(μ, ρ, E, Σ, Δt, T, basis, mesh, γ, α, n_iter, solver)
    return μ, ρ, E, Σ
end
run_forward(μ0, ρ0, E0, Δt, T, basis, mesh, γ, α; n_iter=10, solver=:pcg) =
    run_forward(μ0, ρ0, E0, Δt, T, basis, mesh, γ, α, Int(n_iter), solver)
# forward.jl:137
run_forward_store(μ0, ρ0, E0, Δt, T, basis, mesh, γ, α; n_iter=10, solver=:pcg) =
    IGRHip.run_forward_store(ctx_for(basis, mesh), μ0, ρ0, E0, Δt, T, γ, α; n_iter=Int(n_iter), solver=solver)
# adjoint_pde.jl:466 -> (aμ, aρ, aE)
function run_adjoint_conservative(snapshots::Snapshots, n_steps, Δt, T, basis, mesh, γ, α;
                                  n_iter=10, objective=:l2_density, solver=:pcg)
    a = IGRHip.run_adjoint_conservative(snapshots, n_steps, Δt, T, γ, α; n_iter=Int(n_iter), objective=objective,
                                        solver=solver)
    return ntuple(v -> reshape(a[v], size(mesh.x)), 3)
end
# enzyme_adjoint.jl:267: the IC family at amplitude A on its own mesh, the
# fixed-step objective's gradient, mapped to primitive adjoints (:306-311)
function enzyme_adjoint_ic(A, p, N_e, L, γ, α, CFL, T; n_iter::Int=10, ic_type::Symbol=:acoustic_pulse,
                           solver::Symbol=:pcg, objective::Symbol=:l2_density)
    objective in (:l2_density, :l2_pressure) ||
        error("Unknown objective: $objective (use :l2_density or :l2_pressure)")
    c = ctx_key((:enzyme1d, p, N_e, L), () -> Context(1, p, N_e, 1, L, 1.0))
    n_p = p + 1
    Δx = L / N_e
    max_ws_est = 2.5
    Δt = CFL * Δx / ((2 * p + 1) * max_ws_est)
    n_steps = ceil(Int, T / Δt)
    Δt = T / n_steps
    μ0 = zeros(n_p, N_e); ρ0 = zeros(n_p, N_e); E0 = zeros(n_p, N_e)
    IGRHip.build_ic!(c, (μ0, ρ0, E0), ic_type, A, γ)
    r = IGRHip.enzyme_adjoint_ic_1d(c, μ0, ρ0, E0, n_steps, Δt, γ, α; n_iter=n_iter, solver=solver,
                                    objective=objective)
    return r[1], reshape(r[2], n_p, N_e), reshape(r[3], n_p, N_e), reshape(r[4], n_p, N_e)
end
# sensitivity.jl:475 -> (μf, ρf, Ef, Σ, sμf, sρf, sEf, sΣ)
function run_sensitivity(μ0, ρ0, E0, sμ0, sρ0, sE0, Δt, T, basis, mesh, γ, α;
                         n_iter=10, elliptic_rhs=:form2, solver=:pcg, sγ=0.0, callback=nothing)
    callback === nothing || error("run_sensitivity: callback is not supported on the GPU path")
    c = ctx_for(basis, mesh)
    q, Σ, s, ς = IGRHip.run_sensitivity(c, pack(μ0, ρ0, E0), pack(sμ0, sρ0, sE0), Δt, T, γ, α;
                                        n_iter=Int(n_iter), solver=solver, elliptic_rhs=elliptic_rhs, sγ=sγ)
    sz = size(μ0)
    return (shaped(c, q, μ0)..., reshape(Σ, sz), shaped(c, s, μ0)..., reshape(ς, sz))
end

# initial_conditions.jl:24-210, built on the GPU
function init_entropy_wave(mesh, γ)
    f = ntuple(_ -> zeros(size(mesh.x)), 3)
    IGRHip.build_ic!(ctx_x(mesh.x, mesh.L), f, :init_entropy_wave, 0.0, γ)
    return f
end
function init_constant_state(mesh, γ; ρ0=1.0, u0=1.0, p0=1.0)
    f = ntuple(_ -> zeros(size(mesh.x)), 3)
    IGRHip.build_ic!(ctx_x(mesh.x, mesh.L), f, :constant, 0.0, γ; prm=[ρ0, u0, p0])
    return f
end
build_acoustic_pulse_ic!(μ0, ρ0, E0, A, mesh_x, L, γ) =
    IGRHip.build_ic!(ctx_x(mesh_x, L), (μ0, ρ0, E0), :acoustic_pulse, A, γ)
build_entropy_wave_ic!(μ0, ρ0, E0, A, mesh_x, L, γ) =
    IGRHip.build_ic!(ctx_x(mesh_x, L), (μ0, ρ0, E0), :entropy_wave, A, γ)
build_density_pulse_ic!(μ0, ρ0, E0, A, mesh_x, L, γ) =
    IGRHip.build_ic!(ctx_x(mesh_x, L), (μ0, ρ0, E0), :density_pulse, A, γ)
build_pressure_pulse_ic!(μ0, ρ0, E0, A, mesh_x, L, γ) =
    IGRHip.build_ic!(ctx_x(mesh_x, L), (μ0, ρ0, E0), :pressure_pulse, A, γ)
build_velocity_pulse_ic!(μ0, ρ0, E0, A, mesh_x, L, γ) =
    IGRHip.build_ic!(ctx_x(mesh_x, L), (μ0, ρ0, E0), :velocity_pulse, A, γ)
const IC_BUILDERS = Dict{Symbol, Tuple{Function, String}}(
    :acoustic_pulse => (build_acoustic_pulse_ic!, "ρ=1+A sin, u=A sin, p=1+A sin"),
    :entropy_wave   => (build_entropy_wave_ic!,   "ρ=1+A sin, u=1, p=1"),
    :density_pulse  => (build_density_pulse_ic!,  "ρ=1+A sin, u=0, p=1"),
    :pressure_pulse => (build_pressure_pulse_ic!, "ρ=1, u=0, p=1+A sin"),
    :velocity_pulse => (build_velocity_pulse_ic!, "ρ=1, u=A sin, p=1"),
)
end # module D1

end # module
