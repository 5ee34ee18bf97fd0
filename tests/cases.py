"""Matched set-ups of one configuration on the GPU (libmfrbf context) and on
the CPU oracle, from the same seeded cloud and the same initial state."""
import numpy as np

import oracle as O
from mfrbf import _lib as L
from mfrbf.configs import advection_dt, euler_dt


class Case:
    pass


def build(spec, side=None, use_gpu=True, weights="oracle", hv=None, visc=None, bc=("dirichlet", "nothing"),
          knn="grid", node_order=0, tom=None, order=None, igr=None, igr_maxiter=20, c_rv=1.0, c_uw=1.0):
    """weights: 'oracle' installs the oracle's weights on the GPU (so the
    rhs! comparison isolates the stencil kernels), 'gpu' keeps the GPU's own.
    hv = (k, c): Flyer source; tom = c: Tominec source; igr = alpha: IGR
    source; order: the sources in keyword order, names 'flyer', 'tominec',
    'visc', 'igr' (default: that order); c_rv, c_uw: the viscosity
    source's coefficients."""
    c = Case()
    c.spec = spec
    side = side or spec.side
    x, y, tag, nx, ny = O.cloud_square(side, spec.seed)
    c.x, c.y, c.tag = x, y, tag
    N = len(x)
    c.N, c.k = N, spec.k
    if knn == "brute":
        idx, d2 = O.knn_brute(x, y, spec.k)
    else:
        idx, d2 = O.knn_grid(x, y, spec.k)
    c.idx, c.d2 = idx, d2
    c.dx_min, c.dx_avg = O.dx_stats(d2[:, 1])
    ptr = O.ell_ptr(N, spec.k)
    c.ptr = ptr
    wx, wy, rc = O.weights(x, y, ptr, idx, spec.p, spec.deg, 1)
    assert rc == 0
    c.wx, c.wy = wx.reshape(N, -1), wy.reshape(N, -1)
    euler = spec.eq == "euler"
    c.nvar = 4 if euler else 1
    c.eq = O.EQ_EULER if euler else O.EQ_ADVECTION
    pts = np.c_[x, y]
    from mfrbf.equations import CompressibleEulerEquations2D, LinearScalarAdvectionEquation2D

    eqobj = CompressibleEulerEquations2D(spec.gamma) if euler else LinearScalarAdvectionEquation2D(spec.adv)
    c.u0 = np.ascontiguousarray(np.asarray(spec.ic(pts, 0.0, eqobj), dtype=np.float64).reshape(c.nvar, N))
    c.dt = euler_dt(spec, c.u0, c.dx_min) if euler else advection_dt(spec, c.dx_min)
    # boundary: inflow tag 1, outflow tag 2
    inflow = np.nonzero(tag == 1)[0].astype(np.int32)
    outflow = np.nonzero(tag == 2)[0].astype(np.int32)
    n_in = np.stack([nx[inflow], ny[inflow]], 1)
    n_out = np.stack([nx[outflow], ny[outflow]], 1)
    kinds = {"dirichlet": (L.MF_BC_DIRICHLET, O.BC_DIRICHLET), "nothing": (L.MF_BC_DO_NOTHING, O.BC_NOTHING),
             "slip": (L.MF_BC_SLIP_WALL, O.BC_SLIP)}
    c.bc = [(kinds[bc[0]], inflow, n_in), (kinds[bc[1]], outflow, n_out)]
    params = spec.ic.params()
    fn = spec.ic.code
    c.hv_w = None
    c.hv_gamma = 0.0
    if hv is not None:
        hk, hc = hv
        hx, hy, rc = O.weights(x, y, ptr, idx, spec.p, spec.deg, 2 * hk)
        assert rc == 0
        c.hv_w = (hx + hy).reshape(N, -1)
        c.hv_gamma = hc * c.dx_min ** (2 * hk)
    c.lap_w = None
    c.tom_gamma = 0.0
    if tom is not None:
        lx, ly, rc = O.weights(x, y, ptr, idx, spec.p, spec.deg, 2)
        assert rc == 0
        c.lap_w = (lx + ly).reshape(N, -1)
        c.tom_gamma = tom * c.dx_min ** 4.5
    names = {"flyer": (O.SRC_FLYER, L.MF_SRC_FLYER), "tominec": (O.SRC_TOMINEC, L.MF_SRC_TOMINEC),
             "visc": (O.SRC_VISCOSITY, L.MF_SRC_VISCOSITY), "igr": (O.SRC_IGR, L.MF_SRC_IGR)}
    order = list(order or ["flyer", "tominec", "visc", "igr"])
    hv_first = order.index("flyer") < order.index("visc") if "visc" in order and "flyer" in order else True
    c.visc = visc
    rv_mode = O.RV_NONE
    if visc is not None:
        rv_mode = {"residual": O.RV_RESIDUAL, "upwind": O.RV_UPWIND}[visc]
    c.prob = O.Problem(x, y, ptr, idx, wx, wy, c.eq, c.nvar, gamma=spec.gamma, adv=spec.adv,
                       boundary=[(ko, nodes, nrm, fn, params, None) for ((kg, ko), nodes, nrm) in c.bc],
                       hv=c.hv_w, hv_gamma=c.hv_gamma, rv_mode=rv_mode, c_rv=c_rv, c_uw=c_uw,
                       dx_avg=c.dx_avg, hv_first=hv_first, lap=c.lap_w, tom_gamma=c.tom_gamma,
                       order=[names[n][0] for n in order], igr_alpha=igr or 0.0, igr_maxiter=igr_maxiter)
    if not use_gpu:
        return c
    ctx = L.Context(0)
    c.ctx = ctx
    if euler:
        ctx.set_equations(L.MF_EQ_EULER_2D, spec.gamma)
    else:
        ctx.set_equations(L.MF_EQ_ADVECTION_2D, spec.adv[0], spec.adv[1])
    ctx.set_tuning(L.MF_TUNE_NODE_ORDER, node_order)
    ctx.set_points(x, y)
    ctx.set_stencil(idx)
    if weights == "oracle":
        ctx.set_operator(L.MF_OP_FLUX, c.wx, c.wy)
    else:
        ctx.compute_operator(spec.p, spec.deg, 1, L.MF_OP_FLUX, want=False)
    ptrs = [0]
    for (_, nodes, _) in c.bc:
        ptrs.append(ptrs[-1] + len(nodes))
    ctx.set_boundary(ptrs, [kg for ((kg, ko), _, _) in c.bc], [fn, fn], np.concatenate([params, params]),
                     np.concatenate([b[1] for b in c.bc]), np.concatenate([b[2] for b in c.bc]))
    if hv is not None:
        if weights == "oracle":
            ctx.set_operator(L.MF_OP_HYPERVISCOSITY, c.hv_w)
        else:
            ctx.compute_operator(spec.p, spec.deg, 2 * hv[0], L.MF_OP_HYPERVISCOSITY, want=False)
        ctx.set_source_hyperviscosity(c.hv_gamma)
    if tom is not None:
        if weights == "oracle":
            ctx.set_operator(L.MF_OP_LAPLACIAN, c.lap_w)
        else:
            ctx.compute_operator(spec.p, spec.deg, 2, L.MF_OP_LAPLACIAN, want=False)
        ctx.set_source_hyperviscosity_tominec(c.tom_gamma)
    if visc is not None:
        mode = {"residual": L.MF_VISC_RESIDUAL, "upwind": L.MF_VISC_UPWIND}[visc]
        ctx.set_source_viscosity(mode, c_rv, c_uw, c.dx_avg, getattr(spec, "approx_order", 3), hv_first)
    if igr is not None:
        ctx.set_source_igr(igr, igr_maxiter)
    ctx.set_source_order([names[n][1] for n in order])
    return c


def euler_jvp(c, u, v, gamma=1.4):
    """Forward-mode J v of du = -(Dx F1(u) + Dy F2(u)) for the Euler flux,
    by the chain rule through the textbook flux (v1 = m1/rho, v2 = m2/rho,
    p = (gamma-1)(E - (m1 v1 + m2 v2)/2), F1 = (m1, m1 v1 + p, m2 v1,
    (E + p) v1), F2 likewise): a derivation independent of the oracle's VJP."""
    rho, m1, m2, E = u
    dr, d1, d2, dE = v
    v1, v2 = m1 / rho, m2 / rho
    dv1, dv2 = (d1 - v1 * dr) / rho, (d2 - v2 * dr) / rho
    p = (gamma - 1.0) * (E - 0.5 * (m1 * v1 + m2 * v2))
    dp = (gamma - 1.0) * (dE - 0.5 * (d1 * v1 + m1 * dv1 + d2 * v2 + m2 * dv2))
    dF1 = np.stack([d1, d1 * v1 + m1 * dv1 + dp, d2 * v1 + m2 * dv1, (dE + dp) * v1 + (E + p) * dv1])
    dF2 = np.stack([d2, d1 * v2 + m1 * dv2, d2 * v2 + m2 * dv2 + dp, (dE + dp) * v2 + (E + p) * dv2])
    return -(np.einsum("is,vis->vi", c.wx, dF1[:, c.idx]) + np.einsum("is,vis->vi", c.wy, dF2[:, c.idx]))
