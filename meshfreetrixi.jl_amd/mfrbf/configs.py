"""The benchmark configurations as problem specifications (synthetic data).

  C1 advection: 4,096 nodes (64 x 64, jitter 0.3h), PHS r^3 + degree 2,
     15-point stencils, Gaussian pulse, Dirichlet inflow, SSPRK3, CFL 0.1,
     200 steps.
  C2 Euler + residual viscosity: 262,144 nodes (512 x 512), PHS r^5 +
     degree 3, 30-point stencils, density wave, 100 SSPRK3 steps.
  C3 weights: 1,048,576 ragged stencils (n = 30, 8 % with n in 40..50).
  C4 adjoint: 65,536 nodes (256 x 256), Euler, 100 steps, checkpoints
     every 10, rho0 perturbed by N(0, 1e-3).
  C5 ensemble: 64 Euler clouds of 16k..262k nodes with Flyer
     hyperviscosity (k = 2, c = 1e-4).
Seeds are 20261201 + configuration number.
"""
import numpy as np

from .clouds import boundary_tags, cloud_square, gaussian_noise, ragged_sizes, uniform01
from .equations import DensityWave, GaussianPulse

SEED0 = 20261201


class Spec:
    """Everything needed to set a configuration up on the GPU or the oracle."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def cloud(self):
        x, y, tag, nx, ny = cloud_square(self.side, self.seed)
        return x, y, tag, nx, ny

    def tags(self, tag, nx, ny):
        return boundary_tags(tag, nx, ny)


def c1(side=64):
    return Spec(name="advection", side=side, seed=SEED0 + 1, eq="advection", adv=(1.0, 1.0), gamma=1.4,
                p=3, deg=2, k=15, ic=GaussianPulse((0.3, 0.3), 0.05, (1.0, 1.0)), cfl=0.1, nsteps=200,
                hv=None, rv=None)


def c2(side=512):
    return Spec(name="euler_rv", side=side, seed=SEED0 + 2, eq="euler", adv=(0.0, 0.0), gamma=1.4,
                p=5, deg=3, k=30, ic=DensityWave(1.0, 0.2, (1.0, 1.0), 1.0, 0.0), cfl=0.1, nsteps=100,
                hv=None, rv=dict(c_rv=1.0, c_uw=1.0, polydeg=3), approx_order=3)


def c4(side=256):
    return Spec(name="euler_adjoint", side=side, seed=SEED0 + 4, eq="euler", adv=(0.0, 0.0), gamma=1.4,
                p=5, deg=3, k=30, ic=DensityWave(1.0, 0.2, (1.0, 1.0), 1.0, 0.0), cfl=0.1, nsteps=100,
                hv=None, rv=None, checkpoint_every=10, noise=1e-3, noise_seed=SEED0 + 40)


def c5_sizes(n_clouds=64, seed=SEED0 + 5):
    u = uniform01(seed, 2 * n_clouds)
    n = np.rint(np.exp(np.log(16384.0) + u[0::2] * (np.log(262144.0) - np.log(16384.0))))
    sides = np.rint(np.sqrt(n)).astype(int)
    phases = 2.0 * np.pi * u[1::2]
    return sides, phases


def c5(n_clouds=64):
    """C5 ensemble: Euler + Flyer hyperviscosity (k = 2, c = 1e-4) on clouds
    of 16k..262k nodes (log-uniform), density waves with random phases,
    split-FP32 rhs!, 50 SSPRK33 steps each."""
    sides, phases = c5_sizes(n_clouds)
    return [Spec(name="ensemble_%d" % r, side=int(sides[r]), seed=SEED0 + 500 + r, eq="euler", adv=(0.0, 0.0),
                 gamma=1.4, p=5, deg=3, k=30, ic=DensityWave(1.0, 0.2, (1.0, 1.0), 1.0, float(phases[r])),
                 cfl=0.1, nsteps=50, hv=dict(k=2, c=1e-4), rv=None)
            for r in range(n_clouds)]


def c3_stencils(side=1024, seed=SEED0 + 3):
    """Ragged stencil sizes for the batched-weights configuration."""
    return ragged_sizes(side * side, seed)


def perturbed_density(u0, spec, n):
    u = u0.copy()
    u[0] = u[0] + gaussian_noise(spec.noise_seed, n, spec.noise)
    return u


def euler_dt(spec, u, dx_min):
    """dt = cfl * dx_min / max_i(|v_i| + c_i)."""
    rho, m1, m2, E = u
    v1, v2 = m1 / rho, m2 / rho
    p = (spec.gamma - 1.0) * (E - 0.5 * rho * (v1 * v1 + v2 * v2))
    c = np.sqrt(spec.gamma * p / rho)
    return spec.cfl * dx_min / float(np.max(np.sqrt(v1 * v1 + v2 * v2) + c))


def advection_dt(spec, dx_min):
    return spec.cfl * dx_min / (abs(spec.adv[0]) + abs(spec.adv[1]))
