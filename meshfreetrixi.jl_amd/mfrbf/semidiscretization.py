"""Semidiscretization, rhs!, time integration, callbacks and adjoints.

Mirrors the Trixi entry points MeshfreeTrixi.jl plugs into:
SemidiscretizationHyperbolic + create_cache (pointcloudsolver/rbfsolver.jl:
130-185), semidiscretize, Trixi.rhs! (:397-428), OrdinaryDiffEq's
solve(ode, SSPRK33(stage_limiter!), dt, callbacks) with the
HistoryCallback (callbacks_step/history.jl) and the Zhang-Shu stage
limiter (callbacks_stage/positivity_zhang_shu_point2d.jl).  State arrays
are (nvariables, N) float64 arrays: row v is the reference's
StructArrays.component(u, v).
"""
from collections import OrderedDict

import numpy as np

from . import _lib as L
from .equations import CompressibleEulerEquations2D, LinearScalarAdvectionEquation2D
from .sources import (SourceHyperviscosityFlyer, SourceHyperviscosityTominec, SourceIGR,
                      SourceResidualViscosityTominec, SourceTerms, SourceUpwindViscosityTominec)

PRECISIONS = {"f64": L.MF_PREC_F64, "split32": L.MF_PREC_SPLIT32, "f64_ref": L.MF_PREC_F64_REF}


def _precision(p):
    if isinstance(p, str):
        return PRECISIONS[p]
    return int(p)


class SemidiscretizationHyperbolic:
    """Builds the GPU cache (create_cache, rbfsolver.jl:130-165): uploads the
    cloud and stencils, computes the RBF-FD operators on the device, and
    installs equations, boundary conditions and sources."""

    def __init__(self, domain, equations, initial_condition, solver, boundary_conditions=None,
                 source_terms=None):
        self.domain = domain
        self.equations = equations
        self.initial_condition = initial_condition
        self.solver = solver
        self.boundary_conditions = OrderedDict(boundary_conditions or {})
        if source_terms is None:
            source_terms = SourceTerms()
        elif not isinstance(source_terms, SourceTerms):
            source_terms = SourceTerms(**dict(source_terms))
        self.source_terms = source_terms
        pd = domain.pd
        basis = solver.basis
        ctx = solver.engine.new_context()
        self.ctx = ctx
        if isinstance(equations, CompressibleEulerEquations2D):
            ctx.set_equations(L.MF_EQ_EULER_2D, equations.gamma)
        elif isinstance(equations, LinearScalarAdvectionEquation2D):
            a = equations.advection_velocity
            ctx.set_equations(L.MF_EQ_ADVECTION_2D, a[0], a[1])
        else:
            raise NotImplementedError("equations %r" % type(equations).__name__)
        ctx.set_points(pd.points[:, 0], pd.points[:, 1])
        ctx.set_stencil(pd.neighbors)
        # rbf_differentiation_matrices = compute_flux_operator(solver, domain)
        ctx.compute_operator(basis.phs_power, basis.N, 1, L.MF_OP_FLUX, want=False)
        self._install_boundary()
        self._install_sources()

    # --------------------------------------------------------------
    def _install_boundary(self):
        tags = self.domain.boundary_tags
        ptr, kinds, fns, params, nodes, normals = [0], [], [], [], [], []
        self._host_bc = []  # (entry range, function) evaluated on the host
        for key, bc in self.boundary_conditions.items():
            bd = tags[key]  # KeyError for an unknown tag, as the reference
            kinds.append(bc.kind)
            f = getattr(bc, "device_function", None)
            if bc.kind == L.MF_BC_DIRICHLET and f is None:
                fns.append(L.MF_FN_TABLE)
                params.append(np.zeros(8))
                self._host_bc.append((ptr[-1], ptr[-1] + len(bd.idx), bd.idx, bc.boundary_value_function))
            else:
                fns.append(f.code if f is not None else L.MF_FN_CONST)
                params.append(f.params() if f is not None else np.zeros(8))
            nodes.append(bd.idx)
            normals.append(bd.normals)
            ptr.append(ptr[-1] + len(bd.idx))
        if kinds:
            self.ctx.set_boundary(ptr, kinds, fns, np.concatenate(params), np.concatenate(nodes),
                                  np.concatenate(normals))
        self._nb = ptr[-1]

    def _install_sources(self):
        basis = self.solver.basis
        hv_seen = False
        self.visc_source = None
        self.hv_source = None
        self.tominec_source = None
        self.igr_source = None
        order = []
        for s in self.source_terms.values():
            if isinstance(s, SourceIGR):
                if self.igr_source is not None:
                    raise NotImplementedError("more than one IGR source")
                self.ctx.set_source_igr(s.alpha, s.maxiter)
                self.igr_source = s
                order.append(L.MF_SRC_IGR)
            elif isinstance(s, SourceHyperviscosityTominec):
                if self.tominec_source is not None:
                    raise NotImplementedError("more than one Tominec hyperviscosity source")
                # lap = sum(compute_flux_operator(solver, domain, 2)), hyperviscosity.jl:111-113
                self.ctx.compute_operator(basis.phs_power, basis.N, 2, L.MF_OP_LAPLACIAN, want=False)
                self.ctx.set_source_hyperviscosity_tominec(s.gamma)
                self.tominec_source = s
                order.append(L.MF_SRC_TOMINEC)
            elif isinstance(s, SourceHyperviscosityFlyer):
                if self.hv_source is not None:
                    raise NotImplementedError("more than one Flyer hyperviscosity source")
                self.ctx.compute_operator(basis.phs_power, basis.N, 2 * s.k, L.MF_OP_HYPERVISCOSITY,
                                          want=False)
                self.ctx.set_source_hyperviscosity(s.gamma)
                self.hv_source = s
                hv_seen = True
                order.append(L.MF_SRC_FLYER)
            elif isinstance(s, (SourceResidualViscosityTominec, SourceUpwindViscosityTominec)):
                if self.visc_source is not None:
                    raise NotImplementedError("more than one viscosity source")
                self.ctx.set_source_viscosity(s.mode, s.c_rv, s.c_uw, s.h, s.polydeg, hv_seen)
                self.visc_source = s
                order.append(L.MF_SRC_VISCOSITY)
            else:
                raise NotImplementedError("source %r" % type(s).__name__)
        self.ctx.set_source_order(order)

    def _update_host_bc(self, t):
        if not self._host_bc:
            return
        nv = self.equations.nvariables
        table = np.zeros((self._nb, nv))
        pts = self.domain.pd.points
        for (e0, e1, idx, f) in self._host_bc:
            vals = np.asarray(f(pts[idx], t, self.equations), dtype=np.float64).reshape(nv, -1)
            table[e0:e1] = vals.T
        self.ctx.set_boundary_values(table)

    @property
    def nvariables(self):
        return self.equations.nvariables

    def compute_coefficients(self, t):
        """u[i] = initial_condition(x_i, t) (rbfsolver.jl:174-185)."""
        u = self.initial_condition(self.domain.pd.points, t, self.equations)
        return np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(self.nvariables, -1))


class ODEProblem:
    def __init__(self, u0, tspan, p):
        self.u0, self.tspan, self.p = u0, tuple(tspan), p


def semidiscretize(semi, tspan):
    """Trixi.semidiscretize(semi, tspan) -> ODEProblem."""
    return ODEProblem(semi.compute_coefficients(tspan[0]), tspan, semi)


def rhs_(du, u, semi, t, precision="f64"):
    """Trixi.rhs!(du, u, semi, t): du = rhs(u, t); u is mutated at boundary
    nodes like the reference (rbfsolver.jl:397-428)."""
    semi._update_host_bc(t)
    semi.ctx.rhs_host(t, u, du, _precision(precision))
    return None


# ---------------------------------------------------------- callbacks
class HistoryCallback:
    """HistoryCallback(; approx_order) (history.jl:43-51)."""

    def __init__(self, approx_order):
        self.approx_order = int(approx_order)


class PositivityPreservingLimiterZhangShu:
    """Stage limiter for `variables` ("density"/"pressure") in order
    (Trixi.limiter_zhang_shu!, positivity_zhang_shu_point2d.jl:22-83)."""

    _VAR = {"density": L.MF_VAR_DENSITY, "pressure": L.MF_VAR_PRESSURE}

    def __init__(self, thresholds, variables):
        self.thresholds = tuple(float(t) for t in thresholds)
        self.variables = tuple(self._VAR[v] if isinstance(v, str) else int(v) for v in variables)
        if len(self.thresholds) != len(self.variables) or len(self.thresholds) > 2:
            raise ValueError("one threshold per variable, at most two")


def limiter_zhang_shu_(u, threshold, variable, semi):
    """Trixi.limiter_zhang_shu!(u, threshold, variable, domain, ...) in place."""
    v = PositivityPreservingLimiterZhangShu([threshold], [variable]).variables[0]
    ctx = semi.ctx
    ctx.upload(L.MF_BUF_U, u)
    ctx.limiter_zhang_shu(threshold, v)
    ctx.download(L.MF_BUF_U, u)
    return None


class CallbackSet:
    def __init__(self, *callbacks):
        self.callbacks = [c for c in callbacks if c is not None]


class SSPRK33:
    def __init__(self, stage_limiter=None):
        self.stage_limiter = stage_limiter


class ODESolution:
    def __init__(self, t, u):
        self.t, self.u = t, u


def solve(ode, alg=None, dt=None, callback=None, precision="f64", use_graph=False):
    """solve(ode, SSPRK33(stage_limiter!), dt=dt, callback=CallbackSet(...))
    with fixed steps; the whole integration runs on the device."""
    alg = alg or SSPRK33()
    semi = ode.p
    if semi._host_bc:
        raise NotImplementedError("device time stepping needs device-evaluated Dirichlet functions")
    t0, t1 = ode.tspan
    nsteps = int(round((t1 - t0) / dt))
    if nsteps < 1 or abs(nsteps * dt - (t1 - t0)) > 1e-9 * max(1.0, abs(t1)):
        raise ValueError("tspan must be a whole number of steps dt")
    cbs = callback.callbacks if isinstance(callback, CallbackSet) else ([callback] if callback else [])
    flags = L.MF_STEP_GRAPH if use_graph else 0
    hist = [c for c in cbs if isinstance(c, HistoryCallback)]
    if hist:
        s = semi.visc_source
        if s is not None:
            semi.ctx.set_source_viscosity(s.mode, s.c_rv, s.c_uw, s.h, hist[0].approx_order,
                                          semi.hv_source is not None and _hv_before_visc(semi))
            flags |= L.MF_STEP_HISTORY
    if alg.stage_limiter is not None:
        lim = alg.stage_limiter
        semi.ctx.set_stage_limiter(lim.thresholds, lim.variables)
        flags |= L.MF_STEP_LIMITER
    u = np.ascontiguousarray(ode.u0.copy())
    semi.ctx.upload(L.MF_BUF_U, u)
    semi.ctx.ssprk33(t0, dt, nsteps, _precision(precision), flags)
    semi.ctx.download(L.MF_BUF_U, u)
    return ODESolution([t0, t0 + nsteps * dt], [ode.u0, u])


def _hv_before_visc(semi):
    for s in semi.source_terms.values():
        if isinstance(s, SourceHyperviscosityFlyer):
            return True
        if isinstance(s, (SourceResidualViscosityTominec, SourceUpwindViscosityTominec)):
            return False
    return False


# ------------------------------------------------------------ adjoint
def rhs_vjp(semi, u, lam, t):
    """ubar = (d rhs!/d u)^T lam at state u and time t (flux divergence and
    boundary conditions)."""
    ctx = semi.ctx
    ctx.upload(L.MF_BUF_U, np.ascontiguousarray(u, dtype=np.float64))
    ctx.upload(L.MF_BUF_AUX0, np.ascontiguousarray(lam, dtype=np.float64))
    ctx.rhs_vjp(t)
    return ctx.download(L.MF_BUF_AUX1)


def ssprk33_gradient(semi, u0, t0, dt, nsteps, checkpoint_every=10):
    """L = sum_i rho_i(T)^2 after nsteps SSPRK33 steps from u0 and dL/du0,
    by a reverse sweep with checkpoints every `checkpoint_every` steps."""
    ctx = semi.ctx
    ctx.upload(L.MF_BUF_U, np.ascontiguousarray(u0, dtype=np.float64))
    loss = ctx.ssprk33_vjp(t0, dt, nsteps, checkpoint_every)
    return loss, ctx.download(L.MF_BUF_AUX1)
