# ChainRulesCore rule for IGRHip.D2.loss, the J of the reference's Enzyme
# objectives (IGR_adjoints_2d/enzyme_adjoint.jl:50-76): its pullback is the
# GPU discrete adjoint (igr_gradient), in place of Enzyme reverse mode
# (enzyme_adjoint.jl:270-296), and carries dJ/dγ and dJ/dα as well.
# ChainRulesCore is not a dependency of the reference: include this file
# after IGRHip.jl in an environment that has it.
using ChainRulesCore

function ChainRulesCore.rrule(::typeof(IGRHip.D2.loss), μx0, μy0, ρ0, E0, n_steps, Δt, basis, mesh, γ, α;
                              n_iter::Int=10, solver::Symbol=:pcg, objective::Symbol=:l2_density)
    J, dμx, dμy, dρ, dE, dγ, dα = IGRHip.enzyme_adjoint_ic(IGRHip.ctx_for(basis, mesh), μx0, μy0, ρ0, E0,
                                                           n_steps, Δt, γ, α; n_iter=n_iter, solver=solver,
                                                           objective=objective)
    loss_pullback(J̄) = (NoTangent(), J̄ .* dμx, J̄ .* dμy, J̄ .* dρ, J̄ .* dE,
                        NoTangent(), NoTangent(), NoTangent(), NoTangent(), J̄ * dγ, J̄ * dα)
    return J, loss_pullback
end
