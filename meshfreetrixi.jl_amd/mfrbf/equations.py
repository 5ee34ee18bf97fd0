"""Equations, boundary conditions and analytic initial/boundary functions.

Mirrors Trixi's CompressibleEulerEquations2D / LinearScalarAdvectionEquation2D
(used by MeshfreeTrixi.jl) and equations/PointCloudBCs.jl.  Analytic
functions carry the code of the device-side evaluator (MF_FN_*) so that
time-dependent Dirichlet values are computed on the GPU.
"""
import numpy as np

from . import _lib as L


class LinearScalarAdvectionEquation2D:
    nvariables = 1

    def __init__(self, advection_velocity=(1.0, 1.0)):
        self.advection_velocity = tuple(float(a) for a in advection_velocity)

    def flux(self, u, orientation):
        return self.advection_velocity[orientation] * u


class CompressibleEulerEquations2D:
    nvariables = 4

    def __init__(self, gamma=1.4):
        self.gamma = float(gamma)

    def cons2prim(self, u):
        rho, m1, m2, E = u
        v1, v2 = m1 / rho, m2 / rho
        p = (self.gamma - 1.0) * (E - 0.5 * (m1 * v1 + m2 * v2))
        return np.array([rho, v1, v2, p])

    def prim2cons(self, rho, v1, v2, p):
        return np.array([rho, rho * v1, rho * v2, p / (self.gamma - 1.0) + 0.5 * rho * (v1 * v1 + v2 * v2)])


def nvariables(equations):
    return equations.nvariables


# ----------------------------------------------------------- functions
class AnalyticFunction:
    """f(x, t, equations) with a device-side twin (code MF_FN_*, 8 params)."""

    code = L.MF_FN_CONST

    def params(self):
        return np.zeros(8)

    def __call__(self, x, t, equations):
        raise NotImplementedError


class ConstantState(AnalyticFunction):
    code = L.MF_FN_CONST

    def __init__(self, values):
        self.values = np.asarray(values, dtype=np.float64).ravel()

    def params(self):
        p = np.zeros(8)
        p[:len(self.values)] = self.values
        return p

    def __call__(self, x, t, equations):
        x = np.asarray(x)
        n = x.shape[0] if x.ndim == 2 else 1
        return np.repeat(self.values[:, None], n, axis=1)


class GaussianPulse(AnalyticFunction):
    """exp(-((x - a1 t - x0)^2 + (y - a2 t - y0)^2) / w^2), advected exactly."""

    code = L.MF_FN_PULSE

    def __init__(self, center=(0.3, 0.3), width=0.05, velocity=(1.0, 1.0)):
        self.center, self.width, self.velocity = tuple(center), float(width), tuple(velocity)

    def params(self):
        p = np.zeros(8)
        p[:5] = [self.center[0], self.center[1], self.width, self.velocity[0], self.velocity[1]]
        return p

    def __call__(self, x, t, equations):
        x = np.atleast_2d(x)
        dx = (x[:, 0] - self.velocity[0] * t) - self.center[0]
        dy = (x[:, 1] - self.velocity[1] * t) - self.center[1]
        return np.exp(-(dx * dx + dy * dy) / (self.width * self.width))[None, :]


class DensityWave(AnalyticFunction):
    """Euler density wave rho = r0 + A sin(2 pi (x + y - t (v1 + v2)) + phi),
    constant velocity (v1, v2) and pressure p; an exact solution."""

    code = L.MF_FN_WAVE

    def __init__(self, rho0=1.0, amplitude=0.2, velocity=(1.0, 1.0), pressure=1.0, phase=0.0):
        self.rho0, self.amplitude = float(rho0), float(amplitude)
        self.velocity, self.pressure, self.phase = tuple(velocity), float(pressure), float(phase)

    def params(self):
        p = np.zeros(8)
        p[:6] = [self.rho0, self.amplitude, self.velocity[0], self.velocity[1], self.pressure, self.phase]
        return p

    def __call__(self, x, t, equations):
        x = np.atleast_2d(x)
        v1, v2 = self.velocity
        arg = (x[:, 0] + x[:, 1]) - t * (v1 + v2)
        rho = self.rho0 + self.amplitude * np.sin(2.0 * np.pi * arg + self.phase)
        g = equations.gamma
        E = self.pressure / (g - 1.0) + 0.5 * rho * (v1 * v1 + v2 * v2)
        return np.stack([rho, rho * v1, rho * v2, E])


class LinearGradient(AnalyticFunction):
    """The linear Euler state of the reference's operator tests
    (test/divergence_test.jl:26-36)."""

    code = L.MF_FN_GRADIENT

    def __call__(self, x, t, equations):
        x = np.atleast_2d(x)
        s = 0.1 * x[:, 0]
        q = 0.1 * x[:, 1]
        return np.stack([1.4 + s + q, 4.1 + s + q, 0.0 + s + q, 8.8 + s + q])


def initial_condition_density_wave(**kw):
    return DensityWave(**kw)


# ------------------------------------------------- boundary conditions
class BoundaryConditionDirichlet:
    """u_b <- g(x_b, t), du_b <- 0 (PointCloudBCs.jl:49-63).  An
    AnalyticFunction is evaluated on the device; any other callable
    g(x, t, equations) is evaluated on the host and uploaded per call."""

    kind = L.MF_BC_DIRICHLET

    def __init__(self, boundary_value_function):
        self.boundary_value_function = boundary_value_function

    @property
    def device_function(self):
        f = self.boundary_value_function
        return f if isinstance(f, AnalyticFunction) else None


class BoundaryConditionDoNothing:
    """Leaves u_b and du_b unchanged (PointCloudBCs.jl:108-115)."""

    kind = L.MF_BC_DO_NOTHING


class _SlipWall:
    """Projects the momentum onto the wall and zeroes du momentum
    (PointCloudBCs.jl:87-106)."""

    kind = L.MF_BC_SLIP_WALL

    def __repr__(self):
        return "boundary_condition_slip_wall"


boundary_condition_slip_wall = _SlipWall()
