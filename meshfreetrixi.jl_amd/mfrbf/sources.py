"""Source terms: SourceTerms and the dissipation sources.

Mirrors sources/generic_sources.jl:7-18 and sources/hyperviscosity.jl.
The operators and per-call work run on the GPU; these objects carry the
parameters and are installed into a semidiscretization's context.
"""
from collections import OrderedDict

from . import _lib as L


class SourceTerms:
    """SourceTerms(; kwargs...): sources applied in keyword order
    (generic_sources.jl:7-18, rbfsolver.jl:388-395)."""

    def __init__(self, **sources):
        self.sources = OrderedDict(sources)

    def __iter__(self):
        return iter(self.sources.values())

    def __len__(self):
        return len(self.sources)

    def keys(self):
        return self.sources.keys()

    def values(self):
        return self.sources.values()

    def items(self):
        return self.sources.items()


class SourceHyperviscosityFlyer:
    """du -= gamma * (Dx^2k + Dy^2k) u with gamma = c * dx_min^(2k)
    (hyperviscosity.jl:14-64)."""

    def __init__(self, solver, equations, domain, k=2, c=1.0):
        self.k = int(k)
        self.c = float(c)
        self.gamma = self.c * domain.pd.dx_min ** (2 * self.k)


class SourceResidualViscosityTominec:
    """Residual viscosity: eps = min(eps_rv, eps_uw), du -= sum_d D_d^T(eps D_d u)
    (hyperviscosity.jl:181-244, :382-409)."""

    mode = L.MF_VISC_RESIDUAL

    def __init__(self, solver, equations, domain, c_rv=1.0, c_uw=1.0, polydeg=4):
        self.c_rv, self.c_uw, self.polydeg = float(c_rv), float(c_uw), int(polydeg)
        self.h = domain.pd.dx_avg


class SourceUpwindViscosityTominec:
    """Upwind viscosity: eps = eps_uw (hyperviscosity.jl:148-167, :351-380)."""

    mode = L.MF_VISC_UPWIND

    def __init__(self, solver, equations, domain, c=1.0, c_uw=1.0, polydeg=4):
        self.c_rv, self.c_uw, self.polydeg = float(c), float(c_uw), int(polydeg)
        self.h = domain.pd.dx_avg


class SourceHyperviscosityTominec:
    """du -= gamma * lap' (lap u), lap = Dxx + Dyy from the solver's basis,
    gamma = c * dx_min^4.5 (hyperviscosity.jl:66-134).  The reference forms
    H = lap' * lap as a sparse product; the engine applies its two factors
    (the stencil pass stores lap u, the transposed pass gathers lap')."""

    def __init__(self, solver, equations, domain, c=1.0):
        self.c = float(c)
        self.gamma = self.c * domain.pd.dx_min ** 4.5


class SourceIGR:
    """Information geometric regularization (sources/IGR.jl:17-239): sigma
    from `maxiter` conjugate-gradient iterations on
    (1/rho - alpha (Dx diag(1/rho) Dx + Dy diag(1/rho) Dy)) sigma =
    alpha (trace^2 + trace_sq), then du_m1 -= Dx sigma, du_m2 -= Dy sigma.
    The reference passes linear_solver = cg and calls it as
    linear_solver(sigma, A, b; maxiter = 20) (IGR.jl:189), which
    IterativeSolvers.jl provides as cg!(sigma, A, b); that is what runs here.
    Euler equations; it must be the last source in SourceTerms."""

    def __init__(self, solver, equations, domain, alpha=1.0, linear_solver="cg", maxiter=20):
        if linear_solver not in ("cg", "cg!"):
            raise NotImplementedError("SourceIGR: only the conjugate-gradient solver is provided")
        self.alpha = float(alpha)
        self.maxiter = int(maxiter)
