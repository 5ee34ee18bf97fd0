"""Writes the oracle-generated fixtures under tests/golden/:

c1_advection.npz: the C1 case (linear advection of a Gaussian pulse on
64 x 64 jittered nodes, PHS r^3 + degree-2 polynomials, 15-point stencils,
Dirichlet inflow, do-nothing outflow): the cloud, the kNN stencils, the
Dx/Dy weights, du of rhs! at t = 0, and the state after 200 SSPRK33 steps.

euler_rv_1024.npz: compressible Euler with the residual viscosity and the
history callback (approx_order 3) on 32 x 32 jittered nodes (PHS r^5 +
degree 3, 30-point stencils, density wave): the initial state and the
state after 3 SSPRK33 steps, with eps, its codes and approx_du of the last
rhs!.

They are regression fixtures of the oracle and fixed targets for the GPU
path.  Run from the repository root after building the oracle:
    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "meshfreetrixi.jl_amd")):
    sys.path.insert(0, p)

import cases  # noqa: E402
from mfrbf import configs  # noqa: E402

NSTEPS = 200


def compute_c1():
    c = cases.build(configs.c1(), use_gpu=False)
    du0 = c.prob.rhs(c.u0.copy(), 0.0)[0]
    u = c.u0.copy()
    c.prob.ssprk33(u, 0.0, c.dt, NSTEPS)
    return dict(x=c.x, y=c.y, idx=np.ascontiguousarray(c.idx, dtype=np.int32), wx=c.wx, wy=c.wy, u0=c.u0,
                du0=du0, u_final=u, dt=np.float64(c.dt), nsteps=np.int64(NSTEPS))


RV_STEPS = 3


def compute_euler_rv():
    c = cases.build(configs.c2(), side=32, use_gpu=False, visc="residual")
    u = c.u0.copy()
    c.prob.ssprk33(u, 0.0, c.dt, RV_STEPS, approx_order=3)
    approx_du = c.prob.approx_du.reshape(c.u0.shape).copy()
    # the step's last rhs! (f at the new state, kept for the next step)
    _, eps, epsc = c.prob.rhs(u.copy(), RV_STEPS * c.dt)
    return dict(u0=c.u0, u_final=u, dt=np.float64(c.dt), nsteps=np.int64(RV_STEPS), eps=eps, eps_code=epsc,
                approx_du=approx_du)


FIXTURES = {"c1_advection.npz": compute_c1, "euler_rv_1024.npz": compute_euler_rv}

if __name__ == "__main__":
    for name, fn in FIXTURES.items():
        out = os.path.join(HERE, name)
        np.savez_compressed(out, **fn())
        print("wrote %s (%d bytes)" % (out, os.path.getsize(out)))
