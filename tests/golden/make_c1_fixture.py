"""Writes tests/golden/c1_advection.npz: the C1 case (linear advection of a
Gaussian pulse on 64 x 64 jittered nodes, PHS r^3 + degree-2 polynomials,
15-point stencils, Dirichlet inflow, do-nothing outflow) computed by the C
oracle: the cloud, the kNN stencils, the Dx/Dy weights, du of rhs! at
t = 0, and the state after 200 SSPRK33 steps.  It is a regression fixture
of the oracle and a fixed target for the GPU path.

Run from the repository root after building the oracle:
    python tests/golden/make_c1_fixture.py
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


def compute():
    c = cases.build(configs.c1(), use_gpu=False)
    du0 = c.prob.rhs(c.u0.copy(), 0.0)[0]
    u = c.u0.copy()
    c.prob.ssprk33(u, 0.0, c.dt, NSTEPS)
    return dict(x=c.x, y=c.y, idx=np.ascontiguousarray(c.idx, dtype=np.int32), wx=c.wx, wy=c.wy, u0=c.u0,
                du0=du0, u_final=u, dt=np.float64(c.dt), nsteps=np.int64(NSTEPS))


if __name__ == "__main__":
    out = os.path.join(HERE, "c1_advection.npz")
    np.savez_compressed(out, **compute())
    print("wrote %s (%d bytes)" % (out, os.path.getsize(out)))
