"""GPU parity: libmfrbf (through the C ABI) against the CPU oracle on the same
seeded inputs.  Tolerances: integer outputs bit-exact; FP64 rhs! 1e-12
norm-wise relative (max|d| / max|ref|); split-FP32 rhs! 1e-6."""
import os

import numpy as np
import pytest

import cases
import oracle as O
from conftest import GOLDEN, rel_inf
from mfrbf import _lib as L
from mfrbf import configs

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def _say(msg):
    print(msg, flush=True)


# ----------------------------------------------------------------- kNN
@pytest.mark.parametrize("side,k,seed", [(40, 15, 20261202), (48, 30, 20261203), (33, 50, 5), (30, 64, 20261209)])
def test_knn_bit_exact_vs_brute(mf, side, k, seed):
    _say("knn brute side=%d k=%d" % (side, k))
    x, y, _, _, _ = O.cloud_square(side, seed)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    idx, dist, dmin, davg = ctx.knn(k, want_dist=True)
    ib, d2 = O.knn_brute(x, y, k)
    assert np.array_equal(idx, ib)
    assert np.array_equal(dist, np.sqrt(d2))
    omin, oavg = O.dx_stats(d2[:, 1])
    assert dmin == omin
    assert davg == oavg  # the mean in Julia's pairwise order, caller's node order


def test_knn_bit_exact_large(mf):
    _say("knn grid side=256 k=30")
    x, y, _, _, _ = O.cloud_square(256, 20261205)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    idx, _, _, _ = ctx.knn(30)
    ig, _ = O.knn_grid(x, y, 30)
    assert np.array_equal(idx, ig)


@pytest.mark.parametrize("shape", ["horizontal", "vertical", "nearly_collinear", "elongated"])
def test_knn_degenerate_extents(mf, shape):
    """Clouds whose bounding box is flat (collinear points, along either
    axis), nearly flat, or very elongated: the search grid stays small (at
    most N + 1 cells along the long side; a flat box made the cell count
    overflow before) and the stencils and distances equal brute force, ties
    broken by index."""
    _say("knn degenerate extent: %s" % shape)
    rng = np.random.default_rng(20261210)
    n = 600
    t = rng.random(n)
    if shape == "horizontal":
        x, y = t, np.full(n, 0.25)
    elif shape == "vertical":
        x, y = np.full(n, 0.7), t
    elif shape == "nearly_collinear":
        x, y = t, 0.5 + 1e-13 * rng.standard_normal(n)
    else:
        x, y = 1000.0 * t, rng.random(n)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    idx, dist, dmin, davg = ctx.knn(15, want_dist=True)
    ib, d2 = O.knn_brute(x, y, 15)
    assert np.array_equal(idx, ib)
    assert np.array_equal(dist, np.sqrt(d2))
    omin, oavg = O.dx_stats(d2[:, 1])
    assert dmin == omin and davg == oavg


# ------------------------------------------------------------- weights
@pytest.mark.parametrize("p,deg,k,kd", [(3, 2, 15, 1), (5, 3, 30, 1), (5, 3, 30, 4), (3, 3, 20, 1),
                                        (7, 2, 15, 2), (7, 3, 30, 1), (3, 3, 30, 2), (5, 2, 15, 4)])
def test_weights_vs_oracle(mf, p, deg, k, kd):
    _say("weights p=%d deg=%d k=%d kd=%d" % (p, deg, k, kd))
    x, y, _, _, _ = O.cloud_square(40, 20261202)
    N = len(x)
    idx, _ = O.knn_grid(x, y, k)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    ctx.set_stencil(idx)
    wx, wy = ctx.compute_operator(p, deg, kd, L.MF_OP_FLUX, want=True, k=k)
    ox, oy, rc = O.weights(x, y, O.ell_ptr(N, k), idx, p, deg, kd)
    assert rc == 0
    ex = rel_inf(wx.ravel(), ox)
    ey = rel_inf(wy.ravel(), oy)
    _say("  weights rel diff %.2e %.2e  bit-equal fraction %.3f" %
         (ex, ey, np.mean(wx.ravel() == ox)))
    assert ex < 1e-10 and ey < 1e-10


def test_weights_ragged_vs_oracle(mf):
    _say("ragged weights")
    x, y, _, _, _ = O.cloud_square(48, 20261204)
    N = len(x)
    sizes = configs.ragged_sizes(N, 20261204, p_boundary=0.2)
    big, _ = O.knn_grid(x, y, 50)
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    idx = np.concatenate([big[e, :sizes[e]] for e in range(N)]).astype(np.int32)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    ctx.weights_ragged_set(ptr, idx)
    ctx.weights_ragged_run(5, 3, 1)
    wx, wy = ctx.weights_ragged_get()
    ox, oy, rc = O.weights(x, y, ptr, idx, 5, 3, 1)
    assert rc == 0
    assert rel_inf(wx, ox) < 1e-9 and rel_inf(wy, oy) < 1e-9


# ---------------------------------------------------------------- rhs!
def _rhs_pair(c, prec=L.MF_PREC_F64, t=0.0):
    u = c.u0.copy()
    du = np.zeros_like(u)
    c.ctx.rhs_host(t, u, du, prec)
    uo = c.u0.copy()
    duo, eps, epsc = c.prob.rhs(uo, t)
    return u, du, uo, duo, eps, epsc


def test_rhs_advection_f64(mf):
    _say("rhs advection C1")
    c = cases.build(configs.c1())
    u, du, uo, duo, _, _ = _rhs_pair(c, t=0.05)
    assert rel_inf(du, duo) < TOL64
    assert rel_inf(u, uo) < 1e-15


def test_c1_against_fixture(mf):
    """C1 through the device path (kNN, weights, rhs!, 200 graph-replayed
    SSPRK33 steps) against the committed oracle fixture: stencil indices
    bit for bit; weights 1e-9; du and the final state 1e-10 (the weights
    differ from the fixture's in the last bits)."""
    _say("C1 against the committed fixture")
    ref = np.load(os.path.join(GOLDEN, "c1_advection.npz"))
    spec = configs.c1()
    ctx = L.Context(0)
    ctx.set_points(ref["x"], ref["y"])
    idx, _, _, _ = ctx.knn(spec.k)
    assert np.array_equal(idx, ref["idx"])
    wx, wy = ctx.compute_operator(spec.p, spec.deg, 1, L.MF_OP_FLUX, want=True, k=spec.k)
    assert rel_inf(wx, ref["wx"]) < 1e-9 and rel_inf(wy, ref["wy"]) < 1e-9
    c = cases.build(spec, weights="gpu")  # the same cloud; device weights, oracle stencils
    assert np.array_equal(c.x, ref["x"]) and np.array_equal(c.idx, ref["idx"])
    u = ref["u0"].copy()
    du = np.zeros_like(u)
    c.ctx.rhs_host(0.0, u, du, L.MF_PREC_F64)
    assert rel_inf(du, ref["du0"]) < 1e-10
    c.ctx.upload(L.MF_BUF_U, ref["u0"].copy())
    c.ctx.ssprk33(0.0, float(ref["dt"]), int(ref["nsteps"]), flags=L.MF_STEP_GRAPH)
    ug = c.ctx.download(L.MF_BUF_U)
    _say("  final state rel err %.2e" % rel_inf(ug, ref["u_final"]))
    assert rel_inf(ug, ref["u_final"]) < 1e-10


def test_euler_rv_against_fixture(mf):
    """Three SSPRK33 steps of Euler with the residual viscosity and the
    history callback (32 x 32 nodes) against the committed oracle fixture:
    the state to 1e-12; approx_du and eps of an rhs! at the new state to
    1e-10 and 1e-9, codes exactly.  approx_du is a time difference with
    weights of order 1/dt, which scales the 1e-15 differences of the states
    by about 1e3, and eps's residual |approx_du - du| cancels again: on the
    CPU, noise of 2e-15 on u and 2e-12 on approx_du moves eps by 1.4e-11
    and flips no code."""
    _say("Euler + residual viscosity against the committed fixture")
    ref = np.load(os.path.join(GOLDEN, "euler_rv_1024.npz"))
    c = cases.build(configs.c2(), side=32, visc="residual")
    assert np.array_equal(c.u0, ref["u0"])
    dt, n = float(ref["dt"]), int(ref["nsteps"])
    c.ctx.upload(L.MF_BUF_U, ref["u0"].copy())
    c.ctx.ssprk33(0.0, dt, n, flags=L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    ug = c.ctx.download(L.MF_BUF_U)
    assert rel_inf(ug, ref["u_final"]) < 1e-12
    assert rel_inf(c.ctx.approx_du(), ref["approx_du"]) < 1e-10
    c.ctx.rhs(n * dt, L.MF_PREC_F64_REF)
    geps, gcode = c.ctx.viscosity()
    _say("  approx_du rel err %.2e, eps rel err %.2e" % (rel_inf(c.ctx.approx_du(), ref["approx_du"]),
                                                         rel_inf(geps, ref["eps"])))
    assert np.array_equal(gcode, ref["eps_code"])
    assert rel_inf(geps, ref["eps"]) < 1e-9


def test_rhs_euler_f64(mf):
    _say("rhs euler 64x64 k=30")
    c = cases.build(configs.c2(), side=64)
    u, du, uo, duo, _, _ = _rhs_pair(c, t=0.01)
    assert rel_inf(du, duo) < TOL64
    assert rel_inf(u, uo) < 1e-15


def test_rhs_euler_hv_f64(mf):
    _say("rhs euler + flyer hv")
    c = cases.build(configs.c2(), side=64, hv=(2, 1e-4))
    u, du, uo, duo, _, _ = _rhs_pair(c)
    assert rel_inf(du, duo) < TOL64


@pytest.mark.parametrize("visc", ["upwind", "residual"])
def test_rhs_euler_viscosity_f64(mf, visc):
    _say("rhs euler + %s viscosity" % visc)
    c = cases.build(configs.c2(), side=64, visc=visc)
    if visc == "residual":
        # a history so that eps_rv is active: approx_du from two snapshots
        u1 = c.u0 * 1.0001
        th = np.zeros(4)
        sh = np.zeros(4 * c.u0.size)
        ad = np.zeros(c.u0.size)
        O.history_update(3, 0, 0.0, c.u0.ravel().copy(), th, sh, ad)
        O.history_update(3, 1, 0.001, u1.ravel().copy(), th, sh, ad)
        c.prob.approx_du[:] = ad
        c.prob.cfg.success_iter = 1
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.history_update(0.0, 0)
        c.ctx.upload(L.MF_BUF_U, np.ascontiguousarray(u1))
        c.ctx.history_update(0.001, 1)
        assert rel_inf(c.ctx.approx_du(), ad.reshape(4, -1)) < 1e-12
    u, du, uo, duo, eps, epsc = _rhs_pair(c, t=0.001)
    geps, gcode = c.ctx.viscosity()
    assert rel_inf(geps, eps) < 1e-12
    flips = int(np.sum(gcode != epsc))
    _say("  eps code flips %d" % flips)
    assert flips == 0
    assert rel_inf(du, duo) < TOL64


def _rv_history(c, u1):
    """Two history snapshots (c.u0 at t = 0, u1 at t = 0.001) on both sides,
    so that eps_rv is active."""
    th, sh, ad = np.zeros(4), np.zeros(4 * c.u0.size), np.zeros(c.u0.size)
    O.history_update(3, 0, 0.0, c.u0.ravel().copy(), th, sh, ad)
    O.history_update(3, 1, 0.001, u1.ravel().copy(), th, sh, ad)
    c.prob.approx_du[:] = ad
    c.prob.cfg.success_iter = 1
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.history_update(0.0, 0)
    c.ctx.upload(L.MF_BUF_U, np.ascontiguousarray(u1))
    c.ctx.history_update(0.001, 1)


@pytest.mark.parametrize("shape", ["spike", "spike_last", "tie", "flat"])
def test_rhs_residual_norm_extremes(mf, shape):
    """The lexicographic maximum of |u - mean| (the residual viscosity's
    normalisation) on states built so that it sits at one node (a density
    spike, also in the last, partial block), at two nodes of different
    blocks with equal density (decided by the momentum), or is tied
    everywhere in the first component (constant density).  eps and du
    against the oracle."""
    _say("rhs residual viscosity norm extremes: %s" % shape)
    c = cases.build(configs.c2(), side=72, visc="residual")  # N = 5184: several ranges, a partial last one
    N = c.u0.shape[1]
    u = c.u0.copy()
    i, j = N // 2 + 3, N // 9
    if shape == "spike":
        u[0, i] *= 1.5
    elif shape == "spike_last":
        u[0, N - 2] *= 1.5
    elif shape == "tie":
        u[0, i] = u[0, j] = 1.5 * u[0].max()
        u[1, i] += 0.01
        u[1, j] += 0.02
    else:
        u[0, :] = 1.0
    u[3] = np.maximum(u[3], 0.5 * (u[1] ** 2 + u[2] ** 2) / u[0] + 0.5)  # positive pressure
    c.u0 = np.ascontiguousarray(u)
    _rv_history(c, c.u0 * 1.0001)
    for prec in (L.MF_PREC_F64_REF, L.MF_PREC_F64, L.MF_PREC_SPLIT32):
        uu = c.u0.copy()
        du = np.zeros_like(uu)
        c.ctx.rhs_host(0.001, uu, du, prec)
        duo, eps, epsc = c.prob.rhs(c.u0.copy(), 0.001)
        geps, gcode = c.ctx.viscosity()
        _say("  prec=%d rel err %.2e eps err %.2e" % (prec, rel_inf(du, duo), rel_inf(geps, eps)))
        assert np.array_equal(gcode, epsc)
        assert rel_inf(geps, eps) < (1e-12 if prec != L.MF_PREC_SPLIT32 else 1e-6)
        assert rel_inf(du, duo) < (1e-12 if prec != L.MF_PREC_SPLIT32 else TOL32)


@pytest.mark.parametrize("shape", ["wave", "spike", "dip", "tie", "flat"])
def test_rv_norms_vs_oracle(mf, shape):
    """mean(u) (sum / 4N) and ninf, the lexicographic maximum of |u - mean|,
    against the oracle's pairwise mean and brute-force maximum: at one node
    (a density spike; a density dip, which |rho - mean| never ranks first,
    since the mean is a quarter of the average density), at two nodes of
    equal density that the momentum decides, and with the density flat (the
    momentum decides among all nodes)."""
    _say("residual viscosity norms: %s" % shape)
    c = cases.build(configs.c2(), side=48, visc="residual", bc=("nothing", "nothing"))
    u = c.u0.copy()
    inner = np.nonzero(c.tag == 0)[0]
    i, j = inner[len(inner) // 3], inner[2 * len(inner) // 3]
    if shape == "spike":
        u[0, i] *= 1.5
    elif shape == "dip":
        u[0, i] = 0.25 * u[0].min()
    elif shape == "tie":
        u[0, i] = u[0, j] = 1.5 * u[0].max()
        u[1, i] += 0.01
        u[1, j] += 0.02
    elif shape == "flat":
        u[0, :] = 1.0
    u[3] = np.maximum(u[3], 0.5 * (u[1] ** 2 + u[2] ** 2) / u[0] + 0.5)  # positive pressure
    c.u0 = np.ascontiguousarray(u)
    _rv_history(c, c.u0 * 1.0001)
    mean_o, ninf_o = O.rv_norms(c.u0)
    for prec in (L.MF_PREC_F64_REF, L.MF_PREC_SPLIT32):
        uu, du = c.u0.copy(), np.zeros_like(c.u0)
        c.ctx.rhs_host(0.001, uu, du, prec)
        assert np.array_equal(uu, c.u0)  # no boundary pass changed the state
        mean, ninf = c.ctx.viscosity_norms()
        _say("  prec=%d mean err %.2e, ninf err %.2e" % (prec, np.abs(mean - mean_o).max(),
                                                         np.abs(ninf - ninf_o).max()))
        assert np.abs(mean - mean_o).max() <= 1e-14 * np.abs(mean_o).max()
        assert np.all(np.abs(ninf - ninf_o) <= 1e-13 * np.maximum(1.0, np.abs(mean_o)))
        duo, eps, epsc = c.prob.rhs(c.u0.copy(), 0.001)
        geps, gcode = c.ctx.viscosity()
        assert np.array_equal(gcode, epsc)
        assert rel_inf(geps, eps) < (1e-12 if prec != L.MF_PREC_SPLIT32 else 1e-6)


def test_rhs_split32_error(mf):
    _say("rhs euler split32")
    c = cases.build(configs.c2(), side=96)
    u, du, uo, duo, _, _ = _rhs_pair(c, prec=L.MF_PREC_SPLIT32)
    err = rel_inf(du, duo)
    _say("  split32 rel err %.2e" % err)
    assert err < TOL32


def test_rhs_gpu_weights_vs_oracle(mf):
    _say("rhs euler with GPU weights")
    c = cases.build(configs.c2(), side=64, weights="gpu")
    u, du, uo, duo, _, _ = _rhs_pair(c)
    assert rel_inf(du, duo) < 1e-10


@pytest.mark.parametrize("prec", [L.MF_PREC_F64, L.MF_PREC_SPLIT32, L.MF_PREC_F64_REF])
def test_operator_change_after_rhs(mf, prec):
    """The tile kernels read tile-major copies of the operators: after an
    rhs! has built them, new flux and hyperviscosity operators must be what
    the next rhs! uses, bit for bit as in a context that never saw the old
    ones."""
    _say("operator change after rhs! prec %d" % prec)
    spec = configs.c2()
    rng = np.random.default_rng(5)
    out = []
    for fresh in (False, True):
        c = cases.build(spec, side=48, hv=(2, 1e-4))
        wx2 = c.wx * (1.0 + 0.01 * rng.standard_normal(c.wx.shape))
        wy2 = c.wy * (1.0 + 0.01 * rng.standard_normal(c.wy.shape))
        hv2 = 2.0 * c.hv_w
        rng = np.random.default_rng(5)  # the same new operators both times
        if not fresh:
            u = c.u0.copy()
            du = np.zeros_like(u)
            c.ctx.rhs_host(0.0, u, du, prec)  # builds the tiles and the copies
        c.ctx.set_operator(L.MF_OP_FLUX, np.ascontiguousarray(wx2), np.ascontiguousarray(wy2))
        c.ctx.set_operator(L.MF_OP_HYPERVISCOSITY, np.ascontiguousarray(hv2))
        u = c.u0.copy()
        du = np.zeros_like(u)
        c.ctx.rhs_host(0.0, u, du, prec)
        assert c.ctx.tile_info()["tiles_a"] == 1  # the tile kernels ran
        out.append(du)
    assert np.array_equal(out[0], out[1])


# ------------------------------------------------------------- SSPRK3
def test_ssprk33_advection(mf):
    _say("ssprk33 advection 20 steps")
    c = cases.build(configs.c1())
    n = 20
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.ssprk33(0.0, c.dt, n)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, n)
    assert rel_inf(ug, uo) < 1e-12


def test_ssprk33_euler_rv_history(mf):
    _say("ssprk33 euler + residual viscosity 6 steps, graph")
    c = cases.build(configs.c2(), side=48, visc="residual")
    n = 6
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.ssprk33(0.0, c.dt, n, flags=L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, n, approx_order=3)
    assert rel_inf(ug, uo) < 1e-12


def test_ssprk33_rv_slip_walls(mf):
    """Residual viscosity, history and slip walls: each stage hands the
    next rhs! its first boundary pass (slip projection, Dirichlet values at
    the next stage's time) and sums; 8 steps over two calls (graph, then
    eager resume) against the oracle."""
    _say("ssprk33 euler + residual viscosity + slip walls")
    c = cases.build(configs.c2(), side=48, visc="residual", bc=("dirichlet", "slip"))
    u0 = c.u0.copy()
    u0[1:3] += 0.02 * np.random.default_rng(11).standard_normal(u0[1:3].shape)
    c.ctx.upload(L.MF_BUF_U, u0.copy())
    c.ctx.ssprk33(0.0, c.dt, 5, flags=L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    c.ctx.ssprk33(0.0, c.dt, 3, flags=L.MF_STEP_HISTORY, resume=True)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 8, approx_order=3)
    assert rel_inf(ug, uo) < 1e-12


def test_ssprk33_rv_with_stage_limiter(mf):
    """With a stage limiter between the stages each rhs! does its own
    boundary pass and sums: residual viscosity + history + the Zhang-Shu
    limiter against the oracle; a resume that changes the limiter setting is
    refused (the stored stage was formed for the other sequence)."""
    _say("ssprk33 euler + residual viscosity + stage limiter")
    c = cases.build(configs.c2(), side=40, visc="residual")
    u0 = c.u0.copy()
    u0[0, ::97] = 0.3
    lim = [(0.6, O.VAR_DENSITY), (0.5, O.VAR_PRESSURE)]
    c.ctx.set_stage_limiter([t for t, _ in lim], [L.MF_VAR_DENSITY, L.MF_VAR_PRESSURE])
    c.ctx.upload(L.MF_BUF_U, u0.copy())
    c.ctx.ssprk33(0.0, c.dt, 5, flags=L.MF_STEP_LIMITER | L.MF_STEP_HISTORY)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 5, approx_order=3, limiter=lim, stencil=(O.ell_ptr(c.N, c.k), c.idx))
    assert rel_inf(ug, uo) < 1e-12
    # the limited call's stored stage does not carry the next boundary pass
    with pytest.raises(L.MfError) as e:
        c.ctx.ssprk33(0.0, c.dt, 1, flags=L.MF_STEP_HISTORY, resume=True)
    assert e.value.code == L.MF_ERR_STATE


def test_ssprk33_rv_stage_sums_match_separate_passes(mf):
    """The stage hand-over changes only the order of the density, momentum
    and energy sums (per tile of the transposed kernel instead of the sum
    pass's blocks): the same steps with a stage limiter that never acts
    (thresholds 0) run every rhs! with its own passes and agree to 1e-13."""
    _say("stage sums vs separate passes")
    out = []
    for limited in (False, True):
        c = cases.build(configs.c2(), side=48, visc="residual")
        f = L.MF_STEP_HISTORY | L.MF_STEP_GRAPH
        if limited:
            c.ctx.set_stage_limiter([0.0], [L.MF_VAR_DENSITY])
            f |= L.MF_STEP_LIMITER
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.ssprk33(0.0, c.dt, 6, flags=f)
        out.append(c.ctx.download(L.MF_BUF_U))
    assert rel_inf(out[0], out[1]) < 1e-13


@pytest.mark.parametrize("prec,tol,tol_sums", [(L.MF_PREC_F64_REF, 1e-12, 1e-13), (L.MF_PREC_SPLIT32, TOL32, 1e-9)])
def test_ssprk33_stage_sums_every_source(mf, prec, tol, tol_sums):
    """The stage hand-over with every source the transposed kernel applies
    (Flyer hyperviscosity, the residual viscosity with the history callback,
    then Tominec hyperviscosity), slip walls and Dirichlet inflow, in reference order
    and split-FP32: 6 graph-replayed steps against the oracle, and against
    the same steps with every rhs! running its own boundary and sum passes
    (a stage limiter that never acts)."""
    _say("stage sums, flyer + tominec + residual viscosity, prec=%d" % prec)
    out = []
    for limited in (False, True):
        c = cases.build(configs.c2(), side=45, visc="residual", hv=(2, 1e-4), tom=1.0, bc=("dirichlet", "slip"),
                        order=("flyer", "visc", "tominec"))
        u0 = c.u0.copy()
        u0[1:3] += 0.02 * np.random.default_rng(5).standard_normal(u0[1:3].shape)
        f = L.MF_STEP_HISTORY | L.MF_STEP_GRAPH
        if limited:
            c.ctx.set_stage_limiter([0.0], [L.MF_VAR_DENSITY])
            f |= L.MF_STEP_LIMITER
        c.ctx.upload(L.MF_BUF_U, u0.copy())
        c.ctx.ssprk33(0.0, c.dt, 6, prec, f)
        out.append(c.ctx.download(L.MF_BUF_U))
    uo = u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 6, approx_order=3)
    e_or, e_sums = rel_inf(out[0], uo), rel_inf(out[0], out[1])
    _say("  vs oracle %.2e, vs separate passes %.2e" % (e_or, e_sums))
    assert np.abs(out[0] - u0).max() > 1e-6  # the steps moved the state
    assert e_or < tol and e_sums < tol_sums


# ------------------------------------------------------------- adjoint
def test_rhs_vjp_vs_oracle(mf):
    _say("vjp euler")
    c = cases.build(configs.c2(), side=48)
    lam = np.random.default_rng(1).standard_normal(c.u0.shape)
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.upload(L.MF_BUF_AUX0, lam)
    c.ctx.rhs_vjp(0.0)
    ub = c.ctx.download(L.MF_BUF_AUX1)
    ubo = c.prob.vjp(c.u0.copy(), lam, 0.0)
    assert rel_inf(ub, ubo) < 1e-12


@pytest.mark.parametrize("bc", [("nothing", "nothing"), ("dirichlet", "nothing")])
def test_rhs_vjp_forward_mode(mf, bc):
    """The device VJP satisfies <lam, J v> = <J' lam, v> to 1e-12 with J v
    from the forward-mode chain rule through the textbook Euler flux (no
    oracle in the loop)."""
    _say("vjp euler vs forward mode bc=%s" % (bc,))
    c = cases.build(configs.c2(), side=48, bc=bc)
    rng = np.random.default_rng(11)
    lam = rng.standard_normal(c.u0.shape)
    v = rng.standard_normal(c.u0.shape)
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.upload(L.MF_BUF_AUX0, lam)
    c.ctx.rhs_vjp(0.0)
    ub = c.ctx.download(L.MF_BUF_AUX1)
    ubc = c.u0.copy()
    c.prob.rhs(ubc, 0.0)  # the state after the first boundary pass
    fixed = c.bc[0][1] if bc[0] == "dirichlet" else np.zeros(0, np.int64)
    vv = v.copy()
    vv[:, fixed] = 0.0
    jv = cases.euler_jvp(c, ubc, vv)
    jv[:, fixed] = 0.0
    lhs, rhs = float((lam * jv).sum()), float((ub * v).sum())
    assert abs(lhs - rhs) <= 1e-12 * float(np.abs(lam * jv).sum()), (lhs, rhs)


def test_vjp_tiles_change_no_result(mf):
    """The tiled VJP (lam halo staged in LDS) gives the bits of the global
    gather, for rhs! and through SSPRK33 steps with slip walls."""
    _say("vjp tiles on/off")
    c = cases.build(configs.c2(), side=64, bc=("dirichlet", "slip"))
    lam = np.random.default_rng(2).standard_normal(c.u0.shape)
    res = []
    for tiles in (1, 0):
        c.ctx.set_tuning(L.MF_TUNE_TILES, tiles)
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.upload(L.MF_BUF_AUX0, lam)
        c.ctx.rhs_vjp(0.01)
        a = c.ctx.download(L.MF_BUF_AUX1)
        if tiles:
            assert c.ctx.tile_info()["tiles_b"] == 1
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        loss = c.ctx.ssprk33_vjp(0.0, c.dt, 6, 3)
        res.append((a, loss, c.ctx.download(L.MF_BUF_AUX1)))
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1] and np.array_equal(res[0][2], res[1][2])
    assert rel_inf(res[0][0], c.prob.vjp(c.u0.copy(), lam, 0.01)) < 1e-12


# ------------------------------------------------ reference operation order
@pytest.mark.parametrize("cfg,bc", [("c1", ("nothing", "nothing")), ("c2", ("nothing", "nothing")),
                                    ("c2", ("dirichlet", "nothing"))])
def test_rhs_reference_order_bitwise(mf, cfg, bc):
    _say("rhs f64_ref %s %s" % (cfg, bc))
    spec = getattr(configs, cfg)()
    c = cases.build(spec, side=48, bc=bc)
    u, du, uo, duo, _, _ = _rhs_pair(c, prec=L.MF_PREC_F64_REF, t=0.02)
    if bc[0] == "nothing":
        assert np.array_equal(du, duo)  # same operation order, same bits
    else:
        # device sin/exp of the boundary values may differ in the last bit
        assert rel_inf(du, duo) < 1e-14


def test_rhs_reference_order_hv_and_rv(mf):
    _say("rhs f64_ref euler + hv + upwind viscosity")
    c = cases.build(configs.c2(), side=40, bc=("nothing", "nothing"), hv=(2, 1e-4), visc="upwind")
    u, du, uo, duo, eps, _ = _rhs_pair(c, prec=L.MF_PREC_F64_REF)
    geps, _ = c.ctx.viscosity()
    assert np.array_equal(geps, eps)
    assert rel_inf(du, duo) < 1e-15


def test_rhs_full_size_c2(mf):
    """262,144 nodes (the headline configuration): reference-order FP64, the
    parity mode, within 1e-12 of the oracle; split-FP32 within 1e-6.  The
    fused-multiply-add FP64 mode rounds each product-sum once instead of
    twice, a different operation order from the reference's, and measures
    about 1.0e-12 here (the cancellation in the flux divergence scales it), so
    its bar is 5e-12."""
    _say("rhs full size 512x512 (kNN + weights on the GPU, oracle on the CPU)")
    from mfrbf.clouds import cloud_square
    spec = configs.c2()
    x, y, tag, _, _ = cloud_square(512, spec.seed)
    N = len(x)
    ctx = L.Context(0)
    ctx.set_equations(L.MF_EQ_EULER_2D, 1.4)
    ctx.set_points(x, y)
    idx, _, dmin, davg = ctx.knn(30)
    wx, wy = ctx.compute_operator(5, 3, 1, L.MF_OP_FLUX, want=True, k=30)
    inflow = np.nonzero(tag == 1)[0].astype(np.int32)
    prm = spec.ic.params()
    ctx.set_boundary([0, len(inflow)], [L.MF_BC_DIRICHLET], [L.MF_FN_WAVE], prm, inflow,
                     np.zeros((len(inflow), 2)))
    from mfrbf.equations import CompressibleEulerEquations2D
    u0 = np.ascontiguousarray(spec.ic(np.c_[x, y], 0.0, CompressibleEulerEquations2D(1.4)))
    prob = O.Problem(x, y, O.ell_ptr(N, 30), idx, wx, wy, O.EQ_EULER, 4,
                     boundary=[(O.BC_DIRICHLET, inflow, None, O.FN_WAVE, prm, None)])
    duo, _, _ = prob.rhs(u0.copy(), 0.0)
    errs = {}
    for name, prec in (("f64_ref", L.MF_PREC_F64_REF), ("f64", L.MF_PREC_F64), ("split32", L.MF_PREC_SPLIT32)):
        u = u0.copy()
        du = np.zeros_like(u)
        ctx.rhs_host(0.0, u, du, prec)
        errs[name] = rel_inf(du, duo)
    _say("  full-size rel err vs oracle: %s" % errs)
    assert errs["f64_ref"] < 1e-12
    assert errs["f64"] < 5e-12
    assert errs["split32"] < TOL32


# ------------------------------------------------------- launch variants
def _rhs_residual_case(spec, side, hv=None):
    c = cases.build(spec, side=side, visc="residual", hv=hv)
    u1 = c.u0 * 1.0001
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.history_update(0.0, 0)
    c.ctx.upload(L.MF_BUF_U, np.ascontiguousarray(u1))
    c.ctx.history_update(0.001, 1)
    return c


@pytest.mark.parametrize("prec", [L.MF_PREC_F64, L.MF_PREC_SPLIT32, L.MF_PREC_F64_REF])
def test_rhs_batch_variants_bitwise(mf, prec):
    """Every load-batch variant of the stencil and transposed kernels keeps
    the per-node operation order, so all give the same bits."""
    _say("rhs batch variants prec %d" % prec)
    c = _rhs_residual_case(configs.c2(), 48, hv=(2, 1e-4))
    outs = {}
    for fb in (3, 5, 6, 10):
        for tb in (2, 4, 8):
            c.ctx.set_tuning(L.MF_TUNE_FLUX_BATCH, fb, prec)
            c.ctx.set_tuning(L.MF_TUNE_TRANSPOSE_BATCH, tb, prec)
            u = c.u0.copy()
            du = np.zeros_like(u)
            c.ctx.rhs_host(0.001, u, du, prec)
            outs[(fb, tb)] = du
    ref = outs[(3, 2)]
    for key, du in outs.items():
        assert np.array_equal(du, ref), key


def test_timing_entry_points(mf):
    """mf_time_rhs and mf_time_rhs_cold return positive times and leave the
    state and du of an ordinary rhs! unchanged; n_calls < 1 is rejected."""
    _say("timing entry points")
    c = _rhs_residual_case(configs.c2(), 40)
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.rhs(0.001, L.MF_PREC_F64)
    du0 = c.ctx.download(L.MF_BUF_DU)
    warm, parts = c.ctx.time_rhs(0.001, L.MF_PREC_F64, 5)
    cold = c.ctx.time_rhs_cold(0.001, L.MF_PREC_F64, 3)
    _say("  warm %.4f ms cold %.4f ms" % (warm, cold))
    assert 0.0 < warm < 100.0 and 0.0 < cold < 100.0 and np.all(np.isfinite(parts))
    c.ctx.rhs(0.001, L.MF_PREC_F64)
    assert np.array_equal(c.ctx.download(L.MF_BUF_DU), du0)
    for f in (c.ctx.time_rhs_cold, c.ctx.time_rhs):
        with pytest.raises(L.MfError) as e:
            f(0.001, L.MF_PREC_F64, 0)
        assert e.value.code == L.MF_ERR_ARGUMENT


def test_set_tuning_rejects_bad_values(mf):
    c = cases.build(configs.c1(), side=16)
    with pytest.raises(L.MfError):
        c.ctx.set_tuning(L.MF_TUNE_FLUX_BATCH, 7)
    with pytest.raises(L.MfError):
        c.ctx.set_tuning(L.MF_TUNE_TRANSPOSE_BATCH, 3)
    with pytest.raises(L.MfError):
        c.ctx.set_tuning(9, 4)


def test_api_rejects_bad_input(mf):
    """Argument and call-order errors come back as status codes with a
    message, before any kernel runs on bad data."""
    def code(fn):
        with pytest.raises(L.MfError) as e:
            fn()
        return e.value.code

    x, y, _, _, _ = O.cloud_square(12, 5)
    ctx = L.Context(0)
    assert code(lambda: ctx.knn(8)) == L.MF_ERR_STATE  # no points yet
    assert code(lambda: ctx.set_points(x[:1], y[:1])) == L.MF_ERR_ARGUMENT
    xn = x.copy()
    xn[5] = np.nan
    assert code(lambda: ctx.set_points(xn, y)) == L.MF_ERR_ARGUMENT
    xn[5] = np.inf
    assert code(lambda: ctx.set_points(xn, y)) == L.MF_ERR_ARGUMENT
    ctx.set_points(x[:20], y[:20])
    assert code(lambda: ctx.knn(21)) == L.MF_ERR_ARGUMENT  # k > N
    ctx.set_equations(L.MF_EQ_ADVECTION_2D, 1.0, 1.0)
    ctx.set_points(x, y)
    assert code(lambda: ctx.knn(65)) == L.MF_ERR_ARGUMENT
    assert code(lambda: ctx.compute_operator(5, 3, 1, L.MF_OP_FLUX, want=False)) == L.MF_ERR_STATE
    bad = np.zeros((len(x), 15), np.int32)
    bad[3, 4] = len(x)
    assert code(lambda: ctx.set_stencil(bad)) == L.MF_ERR_ARGUMENT
    ctx.knn(15)
    u = np.ones((1, len(x)))
    assert code(lambda: ctx.rhs_host(0.0, u, np.zeros_like(u))) == L.MF_ERR_STATE  # no flux operator
    assert code(lambda: ctx.set_source_hyperviscosity_tominec(1.0)) == L.MF_ERR_STATE  # no lap operator
    assert code(lambda: ctx.set_source_igr(1.0, 20)) == L.MF_ERR_UNSUPPORTED  # advection
    assert code(lambda: ctx.set_source_order([L.MF_SRC_FLYER, L.MF_SRC_FLYER])) == L.MF_ERR_ARGUMENT
    ctx.compute_operator(5, 3, 1, L.MF_OP_FLUX, want=False)
    assert code(lambda: ctx.rhs_host(0.0, u, np.zeros_like(u), 7)) == L.MF_ERR_ARGUMENT  # precision
    assert code(lambda: ctx.ssprk33(0.0, 0.01, 2, resume=True)) == L.MF_ERR_STATE
    assert code(lambda: ctx.ssprk33(0.0, -0.01, 2)) == L.MF_ERR_ARGUMENT
    assert code(lambda: ctx.compute_operator(5, 9, 1, L.MF_OP_FLUX, want=False)) in (L.MF_ERR_ARGUMENT,
                                                                                    L.MF_ERR_UNSUPPORTED)
    # a decreasing tag_ptr is refused before any entry is read (its last
    # element sizes the arrays that the earlier tags' entries fill)
    nodes = np.arange(6, dtype=np.int32)
    kinds = [L.MF_BC_DIRICHLET, L.MF_BC_DO_NOTHING]
    assert code(lambda: ctx.set_boundary([0, 5, 1], kinds, [0, 0], np.zeros(16), nodes, np.zeros((6, 2)))) \
        == L.MF_ERR_ARGUMENT
    # a value-function code on a tag that is not Dirichlet is ignored (it
    # would not fit the packed entry code): the same du as with code 0
    dus = []
    for fn_other in (0, 100):
        ctx.set_boundary([0, 3, 6], kinds, [L.MF_FN_CONST, fn_other], np.ones(16), nodes, np.zeros((6, 2)))
        ctx.upload(L.MF_BUF_U, u)
        ctx.rhs(0.0)
        dus.append(ctx.download(L.MF_BUF_DU))
    assert np.array_equal(dus[0], dus[1])
    # the context still works after the errors
    ctx.upload(L.MF_BUF_U, u)
    ctx.rhs(0.0)
    assert np.all(np.isfinite(ctx.download(L.MF_BUF_DU)))


@pytest.mark.parametrize("k", [13, 25])
def test_rhs_generic_stencil_size(mf, k):
    """Stencil sizes without a compiled specialisation take the generic
    kernels, whose last load batch is padded (weight 0 on the node itself)."""
    _say("rhs generic stencil size k=%d" % k)
    spec = configs.c2()
    spec.k = k
    spec.p, spec.deg = 3, 2
    c = _rhs_residual_case(spec, 40)
    ad = c.ctx.approx_du()
    c.prob.approx_du[:] = ad.ravel()
    c.prob.cfg.success_iter = 1
    for prec, tol in ((L.MF_PREC_F64, TOL64), (L.MF_PREC_F64_REF, TOL64), (L.MF_PREC_SPLIT32, TOL32)):
        u, du, uo, duo, _, _ = _rhs_pair(c, prec=prec, t=0.001)
        assert rel_inf(du, duo) < tol, prec


@pytest.mark.parametrize("side,k", [(12, 30), (24, 64)])
def test_rhs_small_cloud_and_largest_stencil(mf, side, k):
    """A cloud smaller than one 256-node tile (12 x 12 = 144 nodes, 30-point
    stencils, with the Flyer hyperviscosity) and the largest stencil (64
    nodes, r^3 with degree 2, on 24 x 24): rhs! with the residual viscosity
    and a history in the three precisions against the oracle, then 4
    graph-replayed SSPRK33 steps in reference order."""
    _say("rhs small cloud / largest stencil side=%d k=%d" % (side, k))
    spec = configs.c2()
    spec.k = k
    hv = (2, 1e-4)
    if k == 64:
        spec.p, spec.deg = 3, 2
        hv = None  # the 4th derivatives of r^3 are singular at the centre
    c = _rhs_residual_case(spec, side, hv=hv)
    ad = c.ctx.approx_du()
    c.prob.approx_du[:] = ad.ravel()
    c.prob.cfg.success_iter = 1
    for prec, tol in ((L.MF_PREC_F64, TOL64), (L.MF_PREC_F64_REF, TOL64), (L.MF_PREC_SPLIT32, TOL32)):
        u, du, uo, duo, _, _ = _rhs_pair(c, prec=prec, t=0.001)
        _say("  prec %d: rel err %.2e" % (prec, rel_inf(du, duo)))
        assert rel_inf(du, duo) < tol, prec
    c2 = cases.build(spec, side=side, visc="residual", hv=hv)
    c2.ctx.upload(L.MF_BUF_U, c2.u0.copy())
    c2.ctx.ssprk33(0.0, c2.dt, 4, L.MF_PREC_F64_REF, L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    uo = c2.u0.copy()
    c2.prob.ssprk33(uo, 0.0, c2.dt, 4, approx_order=3)
    assert rel_inf(c2.ctx.download(L.MF_BUF_U), uo) < TOL64


# ------------------------------------------------------------ node order
def test_node_order_changes_no_result(mf):
    """The context stores nodes in Morton order; every result in the
    caller's order must equal the one computed in the caller's own order
    (kNN, weights, rhs! in all precisions, viscosity, step, adjoint)."""
    _say("node order: Morton vs caller order")
    from mfrbf.clouds import cloud_square
    x, y, _, _, _ = cloud_square(64, 20261207)
    res = {}
    for order in (0, 1):
        ctx = L.Context(0)
        ctx.set_tuning(L.MF_TUNE_NODE_ORDER, order)
        ctx.set_points(x, y)
        idx, dist, dmin, davg = ctx.knn(30, want_dist=True)
        perm = ctx.node_order()
        assert sorted(perm.tolist()) == list(range(len(x)))
        if order == 1:
            assert np.array_equal(perm, np.arange(len(x)))
        wx, wy = ctx.compute_operator(5, 3, 1, L.MF_OP_FLUX, want=True, k=30)
        res[order] = dict(idx=idx, dist=dist, dmin=dmin, davg=davg, wx=wx, wy=wy, perm=perm)
    a, b = res[0], res[1]
    assert not np.array_equal(a["perm"], b["perm"])
    for key in ("idx", "dist", "wx", "wy"):
        assert np.array_equal(a[key], b[key]), key
    assert a["dmin"] == b["dmin"]
    assert a["davg"] == b["davg"]  # summed in the caller's order either way
    outs = {}
    for order in (0, 1):
        c = cases.build(configs.c2(), side=48, hv=(2, 1e-4), visc="upwind", bc=("dirichlet", "nothing"),
                        node_order=order)
        o = {}
        for prec in (L.MF_PREC_F64, L.MF_PREC_SPLIT32, L.MF_PREC_F64_REF):
            u = c.u0.copy()
            du = np.zeros_like(u)
            c.ctx.rhs_host(0.01, u, du, prec)
            o[prec] = (u, du)
        o["eps"] = c.ctx.viscosity()
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.ssprk33(0.0, c.dt, 3, flags=L.MF_STEP_GRAPH)
        o["step"] = c.ctx.download(L.MF_BUF_U)
        # the adjoint covers the flux divergence and boundary conditions
        c = cases.build(configs.c2(), side=48, node_order=order)
        lam = np.random.default_rng(2).standard_normal(c.u0.shape)
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.upload(L.MF_BUF_AUX0, lam)
        c.ctx.rhs_vjp(0.0)
        o["vjp"] = c.ctx.download(L.MF_BUF_AUX1)
        outs[order] = o
    for key in outs[0]:
        va, vb = outs[0][key], outs[1][key]
        for p, q in zip(va if isinstance(va, tuple) else (va,), vb if isinstance(vb, tuple) else (vb,)):
            assert np.array_equal(p, q), key


# ------------------------------------------------------------- LDS tiles
@pytest.mark.parametrize("cfg", ["c1", "euler_rv_hv", "euler_rv_generic_k"])
def test_lds_tiles_change_no_result(mf, cfg):
    """Staging the stencil halos in LDS keeps every node's operation order:
    rhs! and SSPRK33 steps give the same bits with and without tiles."""
    _say("LDS tiles on/off %s" % cfg)
    outs = {}
    for tiles in (1, 0):
        if cfg == "c1":
            c = cases.build(configs.c1(), side=48)
        elif cfg == "euler_rv_hv":
            c = cases.build(configs.c2(), side=56, visc="residual", hv=(2, 1e-4))
        else:
            spec = configs.c2()
            spec.k, spec.p, spec.deg = 25, 3, 2
            c = cases.build(spec, side=40, visc="residual")
        c.ctx.set_tuning(L.MF_TUNE_TILES, tiles)
        o = {}
        if c.visc == "residual":
            c.ctx.upload(L.MF_BUF_U, c.u0.copy())
            c.ctx.history_update(0.0, 0)
            c.ctx.upload(L.MF_BUF_U, np.ascontiguousarray(c.u0 * 1.0001))
            c.ctx.history_update(0.001, 1)
        for prec in (L.MF_PREC_F64, L.MF_PREC_SPLIT32, L.MF_PREC_F64_REF):
            u = c.u0.copy()
            du = np.zeros_like(u)
            c.ctx.rhs_host(0.001, u, du, prec)
            o[prec] = du
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.ssprk33(0.0, c.dt, 3, flags=L.MF_STEP_GRAPH | L.MF_STEP_HISTORY)
        o["step"] = c.ctx.download(L.MF_BUF_U)
        outs[tiles] = o
    for key in outs[0]:
        assert np.array_equal(outs[0][key], outs[1][key]), key


def test_state_copy(mf):
    c = cases.build(configs.c2(), side=24)
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.copy(L.MF_BUF_AUX0, L.MF_BUF_U)
    assert np.array_equal(c.ctx.download(L.MF_BUF_AUX0), c.u0)


@pytest.mark.parametrize("seed", range(12))
def test_rhs_random_configuration(mf, seed):
    """rhs! on seeded random combinations against the oracle: cloud side 16
    to 40, basis and stencil from (r^3, deg 2, 15), (r^5, deg 3, 30),
    (r^3, deg 3, 20), (r^5, deg 2, 20); Flyer hyperviscosity or not; no,
    upwind or residual viscosity (with a history); the Tominec source where
    the keyword order allows it; Dirichlet, do-nothing or slip on each of the
    two boundary tags.  Reference order within 1e-12 (bitwise where no
    transcendental and no reduction enters), split-FP32 within 1e-6."""
    rng = np.random.default_rng(7000 + seed)
    spec = configs.c2()
    spec.p, spec.deg, spec.k = [(3, 2, 15), (5, 3, 30), (3, 3, 20), (5, 2, 20)][int(rng.integers(4))]
    side = int(rng.integers(16, 41))
    hv = (2, 1e-4) if rng.random() < 0.5 else None
    visc = [None, "upwind", "residual"][int(rng.integers(3))]
    tom = 1e-3 if (visc != "residual" and rng.random() < 0.4) else None
    bc = tuple(str(b) for b in rng.choice(["dirichlet", "nothing", "slip"], 2))
    _say("random rhs: side %d, p %d deg %d k %d, hv %s, visc %s, tominec %s, bc %s" % (
        side, spec.p, spec.deg, spec.k, hv is not None, visc, tom is not None, bc))
    c = cases.build(spec, side=side, hv=hv, visc=visc, tom=tom, bc=bc)
    if visc == "residual":
        _rv_history(c, c.u0 * 1.0001)
    t = 0.001 if visc == "residual" else 0.02
    for prec, tol in ((L.MF_PREC_F64_REF, TOL64), (L.MF_PREC_SPLIT32, TOL32)):
        u, du, uo, duo, eps, epsc = _rhs_pair(c, prec=prec, t=t)
        _say("  prec %d: du rel err %.2e, u rel err %.2e" % (prec, rel_inf(du, duo), rel_inf(u, uo)))
        assert rel_inf(du, duo) < tol
        assert rel_inf(u, uo) < 1e-14
        if visc is not None:
            geps, gcode = c.ctx.viscosity()
            assert rel_inf(geps, eps) < (1e-12 if prec == L.MF_PREC_F64_REF else 1e-6)
            if prec == L.MF_PREC_F64_REF:
                assert np.array_equal(gcode, epsc)


def test_advection_with_flyer_hyperviscosity(mf):
    """The Flyer source on the scalar advection equation (the reference's
    source acts on any equations; k = 1 here, second derivatives of r^3):
    rhs! in the three precisions and 10 graph-replayed SSPRK33 steps in
    reference order against the oracle."""
    _say("advection + Flyer hyperviscosity")
    c = cases.build(configs.c1(), side=40, hv=(1, 0.01))
    for prec, tol in ((L.MF_PREC_F64, TOL64), (L.MF_PREC_F64_REF, TOL64), (L.MF_PREC_SPLIT32, TOL32)):
        u, du, uo, duo, _, _ = _rhs_pair(c, prec=prec, t=0.05)
        assert rel_inf(du, duo) < tol, prec
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.ssprk33(0.0, c.dt, 10, L.MF_PREC_F64_REF, L.MF_STEP_GRAPH)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 10)
    assert rel_inf(c.ctx.download(L.MF_BUF_U), uo) < TOL64


@pytest.mark.parametrize("visc", ["upwind", "residual"])
def test_rhs_viscosity_nonphysical_states(mf, visc):
    """Nodes with negative pressure and nodes with negative density: the
    upwind viscosity takes the sound speed as 0 there
    (hyperviscosity.jl:246-285), and the residual branch compares with it;
    du, eps and its codes in reference order against the oracle."""
    _say("viscosity on non-physical states: %s" % visc)
    c = cases.build(configs.c2(), side=32, visc=visc)
    if visc == "residual":
        u1 = c.u0 * 1.0001
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.history_update(0.0, 0)
        c.ctx.upload(L.MF_BUF_U, np.ascontiguousarray(u1))
        c.ctx.history_update(0.001, 1)
        c.prob.approx_du[:] = c.ctx.approx_du().ravel()
        c.prob.cfg.success_iter = 1
    u = c.u0.copy()
    rng = np.random.default_rng(21)
    low_p = rng.choice(c.N, 20, replace=False)
    u[3, low_p] = 0.1 * (u[1, low_p] ** 2 + u[2, low_p] ** 2) / u[0, low_p]  # E below the kinetic energy: p < 0
    neg_rho = rng.choice(np.setdiff1d(np.arange(c.N), low_p), 10, replace=False)
    u[0, neg_rho] = -0.5
    c.u0 = u
    ug, du, uo, duo, eps_o, code_o = _rhs_pair(c, prec=L.MF_PREC_F64_REF, t=0.0)
    eps, code = c.ctx.viscosity()
    if visc == "residual":
        assert np.any(code_o == 0) and np.any(code_o == 1)  # both branches decide somewhere
    assert rel_inf(du, duo) < TOL64
    assert rel_inf(eps, eps_o) < 1e-12
    assert np.array_equal(code, code_o)
