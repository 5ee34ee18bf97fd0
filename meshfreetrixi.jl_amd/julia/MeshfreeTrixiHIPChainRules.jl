# ChainRulesCore reverse rule for the rhs! of a PointCloudSolver running on
# RBFFDEngineHIP (no counterpart in the reference).  Load it after
# MeshfreeTrixiHIP and ChainRulesCore:
#
#     using MeshfreeTrixiHIP, ChainRulesCore
#     include(joinpath(pkgdir_of_this_file, "MeshfreeTrixiHIPChainRules.jl"))
#     using .MeshfreeTrixiHIPChainRules
#
# The primal is du = rhs_value(u, t, semi), the semidiscretization's rhs!
# (Trixi.rhs!(du_ode, u_ode, semi, t)) on a copy of u, so the caller's u is
# not mutated at boundary nodes.  Its pullback is mf_rhs_vjp through
# rhs_vjp!: ubar = (d du / d u)' dy, with the boundary passes included (a
# Dirichlet node's state is overwritten by its boundary value, so its ubar
# is zero).  t and semi get no tangent.
module MeshfreeTrixiHIPChainRules

using ChainRulesCore
using Trixi
using ..MeshfreeTrixiHIP: rhs_vjp!

"""
    rhs_value(u, t, semi)

The time derivative du of the state u at time t, without mutating u.
"""
function rhs_value(u, t, semi)
    du = similar(u)
    Trixi.rhs!(du, copy(u), semi, t)
    return du
end

function ChainRulesCore.rrule(::typeof(rhs_value), u, t, semi)
    du = rhs_value(u, t, semi)
    function rhs_value_pullback(dy)
        ubar = similar(u)
        rhs_vjp!(ubar, u, unthunk(dy), t, semi.solver.engine)
        return NoTangent(), ubar, NoTangent(), NoTangent()
    end
    return du, rhs_value_pullback
end

export rhs_value

end # module
