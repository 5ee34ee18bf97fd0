"""
    MeshfreeTrixiHIP

RBF-FD engine for MeshfreeTrixi.jl on AMD MI355X GPUs: `RBFFDEngineHIP`
plugs into the solver's engine slot (src/solvers/rbfsolver.jl:8,33,60-71)
and routes operator construction, `rhs!`, the history callback and the
stage limiter through libmfrbf.so (C ABI in include/mfrbf.h).

Usage (replacing `RBFFDEngine()`):

    using MeshfreeTrixi, MeshfreeTrixiHIP
    solver = PointCloudSolver(basis; engine = RBFFDEngineHIP())
    semi = SemidiscretizationHyperbolic(domain, equations, ic, solver;
                                        boundary_conditions, source_terms)
    sol = solve(semidiscretize(semi, tspan), SSPRK33(); dt, callback)

Written against the header; Julia is not available where this library is
built, so this file is not exercised by its test suite.
"""
module MeshfreeTrixiHIP

using MeshfreeTrixi
using MeshfreeTrixi: AbstractRBFEngine, PointCloudSolver, PointCloudDomain, SourceTerms,
                     SourceHyperviscosityFlyer, SourceResidualViscosityTominec,
                     SourceUpwindViscosityTominec, BoundaryConditionDoNothing, FluxZero
using SparseArrays
using StructArrays
using Trixi

const libmfrbf = get(ENV, "MFRBF_LIB", joinpath(@__DIR__, "..", "mfrbf", "libmfrbf.so"))

# constants of include/mfrbf.h
const MF_OK = Cint(0)
const MF_EQ_ADVECTION_2D, MF_EQ_EULER_2D = Cint(0), Cint(1)
const MF_BC_DIRICHLET, MF_BC_SLIP_WALL, MF_BC_DO_NOTHING = Cint(0), Cint(1), Cint(2)
const MF_FN_TABLE = Cint(4)
const MF_OP_FLUX, MF_OP_HYPERVISCOSITY, MF_OP_LAPLACIAN = Cint(0), Cint(1), Cint(2)
const MF_SRC_FLYER, MF_SRC_TOMINEC, MF_SRC_VISCOSITY, MF_SRC_IGR = Int32(1), Int32(2), Int32(3), Int32(4)
const MF_VISC_NONE, MF_VISC_RESIDUAL, MF_VISC_UPWIND = Cint(0), Cint(1), Cint(2)
const MF_PREC_F64, MF_PREC_SPLIT32, MF_PREC_F64_REF = Cint(0), Cint(1), Cint(2)

const PRECISIONS = Dict(:f64 => MF_PREC_F64_REF, :f64_fma => MF_PREC_F64, :split32 => MF_PREC_SPLIT32)

"""
    RBFFDEngineHIP(; device = 0, precision = :f64)

RBF-FD backend on a HIP device.  `precision = :f64` runs the stencil sums in
the reference's operation order (SparseArrays' CSC `mul!`: ascending column,
rounded products), which reproduces the reference's rhs! bit for bit where
no transcendental enters; `:f64_fma` uses fused multiply-adds in stencil
order (faster by about 1%, about 1e-12 relative from the reference at
262,144 nodes); `:split32` runs FP32 stencil sums on centred differences
(FP64 state, about 5e-7 relative).
"""
mutable struct RBFFDEngineHIP <: AbstractRBFEngine
    ctx::Ptr{Cvoid}
    device::Cint
    precision::Cint
    installed::Bool   # boundary conditions and sources uploaded
    function RBFFDEngineHIP(; device = 0, precision = :f64)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:mf_ctx_create, libmfrbf), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
        rc == MF_OK || error("mf_ctx_create failed ($rc)")
        haskey(PRECISIONS, precision) || error("RBFFDEngineHIP: precision must be :f64, :f64_fma or :split32")
        e = new(h[], Cint(device), PRECISIONS[precision], false)
        finalizer(e -> ccall((:mf_ctx_destroy, libmfrbf), Cint, (Ptr{Cvoid},), e.ctx), e)
        return e
    end
end

function check(e::RBFFDEngineHIP, rc)
    rc == MF_OK && return nothing
    msg = unsafe_string(ccall((:mf_last_error, libmfrbf), Cstring, (Ptr{Cvoid},), e.ctx))
    throw(ErrorException("libmfrbf: " * msg))
end

fieldptrs(u::StructArray) = [pointer(StructArrays.component(u, v))
                             for v in 1:length(StructArrays.components(u))]

# --- create_cache: the GPU holds cloud, stencils and operators
# (replaces pointcloudsolver/rbfsolver.jl:130-165)
function Trixi.create_cache(domain::PointCloudDomain{2}, equations,
                            solver::PointCloudSolver{2, <:Any, <:Any, RBFFDEngineHIP},
                            RealT, uEltype)
    e = solver.engine
    pd = domain.pd
    x = [p[1] for p in pd.points]
    y = [p[2] for p in pd.points]
    if equations isa CompressibleEulerEquations2D
        check(e, ccall((:mf_set_equations, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble),
                       e.ctx, MF_EQ_EULER_2D, equations.gamma, 0.0))
    else
        a = equations.advection_velocity
        check(e, ccall((:mf_set_equations, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble),
                       e.ctx, MF_EQ_ADVECTION_2D, a[1], a[2]))
    end
    check(e, ccall((:mf_set_points, libmfrbf), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Int64),
                   e.ctx, x, y, pd.num_points))
    nv = pd.num_neighbors
    idx = Matrix{Int32}(undef, nv, pd.num_points)   # column per node = row-major per node
    for i in 1:pd.num_points
        idx[:, i] .= pd.neighbors[i] .- 1
    end
    check(e, ccall((:mf_set_stencil, libmfrbf), Cint, (Ptr{Cvoid}, Ptr{Int32}, Int64, Cint),
                   e.ctx, idx, pd.num_points, nv))
    basis = solver.basis
    check(e, ccall((:mf_compute_operator, libmfrbf), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}),
                   e.ctx, basis.approximation_type.Nrbf, basis.N, 1, MF_OP_FLUX, C_NULL, C_NULL))
    nvars = nvariables(equations)
    u_values = StructArray{SVector{nvars, uEltype}}(ntuple(_ -> zeros(uEltype, pd.num_points), nvars))
    return (; pd, engine = e, u_values)
end

# --- boundary conditions and sources are installed once per semidiscretization
function install!(e::RBFFDEngineHIP, domain, equations, boundary_conditions, source_terms, solver)
    tag_ptr = Int32[0]; kinds = Int32[]; fns = Int32[]; params = Float64[]
    nodes = Int32[]; normals = Float64[]
    for (key, bc) in zip(keys(boundary_conditions), boundary_conditions)
        bd = domain.boundary_tags[key]
        # the kinds calc_single_boundary_flux! can apply (PointCloudBCs.jl:49-115);
        # anything else has no method there and raises here
        if bc isa BoundaryConditionDirichlet
            push!(kinds, MF_BC_DIRICHLET)
        elseif bc isa BoundaryConditionDoNothing
            push!(kinds, MF_BC_DO_NOTHING)
        elseif bc === Trixi.boundary_condition_slip_wall
            push!(kinds, MF_BC_SLIP_WALL)
        else
            throw(ArgumentError("RBFFDEngineHIP: unsupported boundary condition $(typeof(bc)) on :$key"))
        end
        push!(fns, MF_FN_TABLE)          # values uploaded per call (any Julia function)
        append!(params, zeros(8))
        append!(nodes, Int32.(bd.idx .- 1))
        for n in bd.normals
            append!(normals, (n[1], n[2]))
        end
        push!(tag_ptr, length(nodes))
    end
    check(e, ccall((:mf_set_boundary, libmfrbf), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Cdouble}, Ptr{Int32}, Ptr{Cdouble}),
                   e.ctx, length(kinds), tag_ptr, kinds, fns, params, nodes, normals))
    hv_first = false
    order = Int32[]  # the SourceTerms keyword order (generic_sources.jl:7-18)
    for s in values(source_terms)
        if s isa SourceIGR
            # sigma by cg! with maxiter = 20 (IGR.jl:189)
            check(e, ccall((:mf_set_source_igr, libmfrbf), Cint, (Ptr{Cvoid}, Cdouble, Cint),
                           e.ctx, s.cache.alpha, 20))
            push!(order, MF_SRC_IGR)
        elseif s isa SourceHyperviscosityTominec
            # lap = sum(compute_flux_operator(solver, domain, 2)) (hyperviscosity.jl:111-113)
            basis = solver.basis
            check(e, ccall((:mf_compute_operator, libmfrbf), Cint,
                           (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}),
                           e.ctx, basis.approximation_type.Nrbf, basis.N, 2, MF_OP_LAPLACIAN,
                           C_NULL, C_NULL))
            check(e, ccall((:mf_set_source_hyperviscosity_tominec, libmfrbf), Cint, (Ptr{Cvoid}, Cdouble),
                           e.ctx, s.cache.gamma))
            push!(order, MF_SRC_TOMINEC)
        elseif s isa SourceHyperviscosityFlyer
            # the cache holds the assembled operator sum(compute_flux_operator(solver,
            # domain, 2k)) and gamma (hyperviscosity.jl:33-50), not k: upload its
            # entries in stencil order
            H = stencil_values(s.cache.hv_differentiation_matrix, domain.pd)
            check(e, ccall((:mf_set_operator, libmfrbf), Cint,
                           (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}),
                           e.ctx, MF_OP_HYPERVISCOSITY, H, C_NULL))
            check(e, ccall((:mf_set_source_hyperviscosity, libmfrbf), Cint, (Ptr{Cvoid}, Cdouble),
                           e.ctx, s.cache.gamma))
            hv_first = true
            push!(order, MF_SRC_FLYER)
        elseif s isa SourceResidualViscosityTominec || s isa SourceUpwindViscosityTominec
            mode = s isa SourceResidualViscosityTominec ? MF_VISC_RESIDUAL : MF_VISC_UPWIND
            P = length(s.cache.time_history) - 1
            check(e, ccall((:mf_set_source_viscosity, libmfrbf), Cint,
                           (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cint, Cint),
                           e.ctx, mode, s.cache.c_rv, s.cache.c_uw, domain.pd.dx_avg, P, hv_first))
            push!(order, MF_SRC_VISCOSITY)
        end
    end
    check(e, ccall((:mf_set_source_order, libmfrbf), Cint, (Ptr{Cvoid}, Ptr{Int32}, Cint),
                   e.ctx, order, length(order)))
end

# --- rhs! (replaces pointcloudsolver/rbfsolver.jl:397-428)
function Trixi.rhs!(du, u, t, domain, equations, initial_condition, boundary_conditions,
                    source_terms, solver::PointCloudSolver{2, <:Any, <:Any, RBFFDEngineHIP}, cache)
    e = solver.engine
    if !e.installed  # Trixi's create_cache does not see the BCs and sources
        install!(e, domain, equations, boundary_conditions, source_terms, solver)
        e.installed = true
    end
    # Dirichlet values g(x_b, t) from the Julia functions
    vals = Float64[]
    for (key, bc) in zip(keys(boundary_conditions), boundary_conditions)
        bd = domain.boundary_tags[key]
        for b in bd.idx
            if bc isa BoundaryConditionDirichlet
                append!(vals, bc.boundary_value_function(domain.pd.points[b], t, equations))
            else
                append!(vals, zeros(nvariables(equations)))
            end
        end
    end
    isempty(vals) || check(e, ccall((:mf_set_boundary_values, libmfrbf), Cint,
                                    (Ptr{Cvoid}, Ptr{Cdouble}), e.ctx, vals))
    up = fieldptrs(u)
    dup = fieldptrs(du)
    check(e, ccall((:mf_rhs_host, libmfrbf), Cint,
                   (Ptr{Cvoid}, Cdouble, Cint, Ptr{Ptr{Cdouble}}, Ptr{Ptr{Cdouble}}),
                   e.ctx, t, e.precision, up, dup))
    return nothing
end

# Entries of an assembled N x N operator (SparseMatrixCSC, the reference's
# compute_flux_operator output, compute_operators.jl:443-452) in stencil
# order: row i, slot j holds A[i, neighbors[i][j]]; an (nv, N) matrix whose
# column i is node i's row, the layout mf_set_operator reads.
function stencil_values(A, pd)
    At = sparse(transpose(A))  # column i of At = row i of A
    cols, vals = rowvals(At), nonzeros(At)
    W = zeros(Float64, pd.num_neighbors, pd.num_points)
    for i in 1:pd.num_points
        for (j, q) in enumerate(pd.neighbors[i])
            for r in nzrange(At, i)
                if cols[r] == q
                    W[j, i] = vals[r]
                    break
                end
            end
        end
    end
    return W
end

# --- HistoryCallback hook: the reference's update_history! loops over the
# sources and calls modify_cache! on each (callbacks_step/history.jl:77-103).
# For a semidiscretization on this engine the history and approx_du live on
# the device, so this more specific method takes the place of that loop for
# the HIP solver only; the reference's modify_cache! methods, and every
# other engine, stay untouched.
const HIPSemi = Trixi.SemidiscretizationHyperbolic{<:Any, <:Any, <:Any, <:Any, <:Any,
                                                   <:PointCloudSolver{2, <:Any, <:Any, RBFFDEngineHIP}}

function MeshfreeTrixi.update_history!(semi::HIPSemi, u, t, approx_order, integrator)
    e = semi.solver.engine
    for source in values(semi.source_terms)
        source isa SourceResidualViscosityTominec || continue
        source.cache.success_iter .= integrator.success_iter  # as modify_cache!, history.jl:96
        check(e, ccall((:mf_state_upload, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cdouble}}),
                       e.ctx, 0, fieldptrs(u)))
        check(e, ccall((:mf_history_update, libmfrbf), Cint, (Ptr{Cvoid}, Cdouble, Cint),
                       e.ctx, t, integrator.success_iter))
    end
    return nothing
end

# --- stage limiter (replaces callbacks_stage/positivity_zhang_shu_point2d.jl:22-83)
function Trixi.limiter_zhang_shu!(u, threshold::Real, variable, domain::PointCloudDomain{2}, equations,
                                  solver::PointCloudSolver{2, <:Any, <:Any, RBFFDEngineHIP}, cache)
    e = solver.engine
    var = variable === Trixi.density ? Cint(0) : Cint(1)
    check(e, ccall((:mf_state_upload, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cdouble}}),
                   e.ctx, 0, fieldptrs(u)))
    check(e, ccall((:mf_limiter_zhang_shu, libmfrbf), Cint, (Ptr{Cvoid}, Cdouble, Cint), e.ctx, threshold, var))
    check(e, ccall((:mf_state_download, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cdouble}}),
                   e.ctx, 0, fieldptrs(u)))
    return nothing
end

# --- adjoint (no counterpart in the reference): ubar = J' lambda of rhs!
# through mf_rhs_vjp (lambda in buffer MF_BUF_AUX0, ubar read from
# MF_BUF_AUX1).  The ChainRulesCore rrule built on it is in
# MeshfreeTrixiHIPChainRules.jl beside this file.
function rhs_vjp!(ubar, u, lambda, t, e::RBFFDEngineHIP)
    check(e, ccall((:mf_state_upload, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cdouble}}),
                   e.ctx, 0, fieldptrs(u)))
    check(e, ccall((:mf_state_upload, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cdouble}}),
                   e.ctx, 2, fieldptrs(lambda)))
    check(e, ccall((:mf_rhs_vjp, libmfrbf), Cint, (Ptr{Cvoid}, Cdouble), e.ctx, t))
    check(e, ccall((:mf_state_download, libmfrbf), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cdouble}}),
                   e.ctx, 3, fieldptrs(ubar)))
    return ubar
end

export RBFFDEngineHIP, rhs_vjp!

end # module
