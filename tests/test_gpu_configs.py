"""The benchmark configurations at their stated sizes, GPU (through the C ABI)
against the CPU oracle on the same seeded inputs.

  C2  kNN (512 x 512, k = 30) bit for bit; Euler + residual viscosity +
      history callback at full size: 10 SSPRK33 steps and the rhs! after
      them in the reference's operation order within 1e-12; 100 steps on
      128 x 128 nodes within 1e-12.
  C3  kNN (1024 x 1024, k = 50, the ragged stencils' prefixes) bit for bit.
  C5  the ensemble paths (contexts stepped concurrently from captured
      graphs; one grouped context whose launches cover all clouds), split-
      FP32, Flyer hyperviscosity, 50 steps, on the smallest, the largest and
      two more of the 64 clouds: each final state within 1e-6 of the
      oracle's FP64 steps and close to the exact travelling wave.
"""
import numpy as np
import pytest

import cases
import oracle as O
from conftest import rel_inf
from mfrbf import _lib as L
from mfrbf import configs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("side,k,seed", [(512, 30, configs.SEED0 + 2), (1024, 50, configs.SEED0 + 3)])
def test_knn_full_size_vs_oracle(mf, side, k, seed):
    """C2's and C3's clouds: the GPU's stencils equal the oracle's grid kNN
    bit for bit, and so do the distances, dx_min and dx_avg (the mean in
    Julia's pairwise order over the caller's node order)."""
    _say("knn full size side=%d k=%d" % (side, k))
    x, y, _, _, _ = O.cloud_square(side, seed)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    idx, dist, dmin, davg = ctx.knn(k, want_dist=True)
    ig, d2 = O.knn_grid(x, y, k)
    assert np.array_equal(idx, ig)
    assert np.array_equal(dist, np.sqrt(d2))
    omin, oavg = O.dx_stats(d2[:, 1])
    assert dmin == omin
    assert davg == oavg


def test_c2_full_size_rv_history_steps(mf):
    """C2 as stated: 262,144 nodes, Euler + residual viscosity (c_rv = c_uw = 1)
    + HistoryCallback(approx_order = 3), reference-order FP64.  10 SSPRK33
    steps from captured graphs within 1e-12 of the oracle's steps; then one
    rhs! on the GPU's state with its history (success_iter = 10, so the
    residual branch is live): du within 1e-12, eps within 1e-10, the choice
    codes equal away from ties (on the smooth density wave eps_rv < eps_uw
    at every node)."""
    _say("C2 full size: residual viscosity + history, 10 steps")
    c = cases.build(configs.c2(), visc="residual")
    n = 10
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.ssprk33(0.0, c.dt, n, L.MF_PREC_F64_REF, L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, n, approx_order=3)
    err = rel_inf(ug, uo)
    _say("  10 steps: state rel err %.2e" % err)
    assert err < 1e-12
    # one rhs! at the GPU's state and history
    ad = c.ctx.approx_du()
    t = n * c.dt
    c.ctx.upload(L.MF_BUF_U, ug)
    c.ctx.rhs(t, L.MF_PREC_F64_REF)
    du = c.ctx.download(L.MF_BUF_DU)
    eps, code = c.ctx.viscosity()
    duo, epso, codeo = c.prob.rhs(ug.copy(), t, success_iter=n, approx_du=ad)
    e_du, e_eps = rel_inf(du, duo), rel_inf(eps, epso)
    _say("  rhs!: du %.2e, eps %.2e, residual branch at %.0f%% of nodes" % (e_du, e_eps, 100 * np.mean(codeo == 0)))
    assert np.mean(codeo == 0) > 0.5  # the residual branch decides (eps_rv < eps_uw)
    assert e_du < 1e-12 and e_eps < 1e-10
    assert np.mean(code != codeo) < 1e-4


def test_c2_rv_history_100_steps_128(mf):
    """C2's 100 steps with the residual viscosity and the history callback,
    on 128 x 128 nodes: reference-order FP64 within 1e-12 of the oracle."""
    _say("C2 100 steps at 128x128: residual viscosity + history")
    c = cases.build(configs.c2(), side=128, visc="residual")
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.ssprk33(0.0, c.dt, 100, L.MF_PREC_F64_REF, L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 100, approx_order=3)
    err = rel_inf(ug, uo)
    _say("  rel err %.2e" % err)
    assert err < 1e-12


def _c5_pick():
    sides, _ = configs.c5_sizes()
    order = np.argsort(sides)
    pick = [int(order[0]), int(order[-1])]
    pick += [r for r in (int(order[len(order) // 3]), int(order[2 * len(order) // 3])) if r not in pick]
    return pick


@pytest.mark.timeout(600)
def test_c5_ensemble_vs_oracle(mf):
    """C5's ensemble path on four of its clouds (the smallest, the largest and
    two between): Euler + Flyer hyperviscosity (k = 2, c = 1e-4), split-FP32
    rhs!, 50 SSPRK33 steps replayed from each context's captured graph with
    all contexts enqueued before any is synchronized.  Each final state is
    within 1e-6 of the oracle's FP64 steps, and the density within 1e-5 of
    the exact travelling wave."""
    from mfrbf.equations import CompressibleEulerEquations2D

    specs = configs.c5()
    pick = _c5_pick()
    cs = []
    for r in pick:
        _say("C5 cloud %d: %d x %d nodes" % (r, specs[r].side, specs[r].side))
        cs.append(cases.build(specs[r], hv=(2, 1e-4), weights="gpu"))
    steps = specs[0].nsteps
    for c in cs:
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    for c in cs:
        c.ctx.ssprk33(0.0, c.dt, steps, L.MF_PREC_SPLIT32, L.MF_STEP_GRAPH | L.MF_STEP_NOSYNC)
    for c in cs:
        c.ctx.synchronize()
    eq = CompressibleEulerEquations2D(1.4)
    for r, c in zip(pick, cs):
        ug = c.ctx.download(L.MF_BUF_U)
        uo = c.u0.copy()
        c.prob.ssprk33(uo, 0.0, c.dt, steps)
        ex = specs[r].ic(np.c_[c.x, c.y], steps * c.dt, eq)
        err, err_ex = rel_inf(ug, uo), float(np.abs(ug[0] - ex[0]).max())
        _say("  cloud %d (N=%d): vs oracle %.2e, density vs exact %.2e" % (r, c.N, err, err_ex))
        assert err < 1e-6
        assert err_ex < 1e-5


@pytest.mark.timeout(600)
def test_c5_grouped_ensemble_vs_oracle(mf):
    """C5 as one grouped context (mf_set_points_grouped): four of its clouds
    (the smallest, the largest and two between) concatenated, each with its
    own stencils, boundary tags and Flyer coefficient, all advanced by every
    launch of one captured SSPRK33 step (split-FP32, 50 steps, the common
    time step).  Each cloud's final state is within 1e-6 of the oracle's
    FP64 steps of that cloud alone, and its density within 1e-5 of the
    exact travelling wave."""
    from mfrbf.ensemble import Ensemble
    from mfrbf.equations import CompressibleEulerEquations2D

    specs = configs.c5()
    pick = _c5_pick()
    _say("grouped ensemble of clouds %s" % pick)
    ens = Ensemble([specs[r] for r in pick])
    steps = specs[0].nsteps
    ens.ctx.upload(L.MF_BUF_U, ens.u0)
    ens.ctx.ssprk33(0.0, ens.dt, steps, L.MF_PREC_SPLIT32, L.MF_STEP_GRAPH)
    ug = ens.ctx.download(L.MF_BUF_U)
    eq = CompressibleEulerEquations2D(1.4)
    for q, r in enumerate(pick):
        c = cases.build(specs[r], hv=(2, 1e-4), use_gpu=False)
        uo = c.u0.copy()
        c.prob.ssprk33(uo, 0.0, ens.dt, steps)
        gr = ens.cloud(q, ug)
        ex = specs[r].ic(np.c_[c.x, c.y], steps * ens.dt, eq)
        err, err_ex = rel_inf(gr, uo), float(np.abs(gr[0] - ex[0]).max())
        _say("  cloud %d (N=%d): vs oracle %.2e, density vs exact %.2e" % (r, c.N, err, err_ex))
        assert err < 1e-6
        assert err_ex < 1e-5


def test_grouped_set_refuses_global_reductions(mf):
    """The residual viscosity reduces over all points, which would couple the
    clouds of a grouped set: rhs! refuses it there.  So do a kNN search over
    the whole set and a stencil that reaches into another cloud."""
    from mfrbf.ensemble import Ensemble

    specs = configs.c5(2)
    for s in specs:
        s.side = 24
    ens = Ensemble(specs)
    with pytest.raises(L.MfError) as e:
        ens.ctx.knn(specs[0].k)
    assert e.value.code == L.MF_ERR_UNSUPPORTED
    bad = np.tile(np.arange(specs[0].k, dtype=np.int32), (ens.N, 1))  # every stencil in cloud 0
    with pytest.raises(L.MfError) as e:
        ens.ctx.set_stencil(bad)
    assert e.value.code == L.MF_ERR_ARGUMENT
    ens.ctx.set_source_viscosity(L.MF_VISC_RESIDUAL, 1.0, 1.0, ens.dx_avg[0], 3, True)
    ens.ctx.upload(L.MF_BUF_U, ens.u0)
    with pytest.raises(L.MfError) as e:
        ens.ctx.rhs(0.0, L.MF_PREC_F64)
    assert e.value.code == L.MF_ERR_UNSUPPORTED


def test_c5_lanes_equal_one_grouped_context(mf):
    """EnsembleLanes (the clouds packed into grouped contexts of at most
    100,000 nodes, each on its own stream) against one grouped context of
    the same clouds: 5 graph-replayed SSPRK33 steps at the common time
    step, split-FP32 with Flyer hyperviscosity, bit for bit."""
    from mfrbf.ensemble import Ensemble, EnsembleLanes
    _say("C5 lanes vs one grouped context")
    specs = configs.c5()[:6]
    g = Ensemble(specs)
    g.ctx.upload(L.MF_BUF_U, g.u0)
    g.ctx.ssprk33(0.0, g.dt, 5, L.MF_PREC_SPLIT32, L.MF_STEP_GRAPH)
    ug = g.ctx.download(L.MF_BUF_U)
    lanes = EnsembleLanes(specs, lane_nodes=100000)
    assert len(lanes.lanes) > 1 and lanes.dt == g.dt
    lanes.upload_initial()
    lanes.ssprk33(0.0, 5)
    assert np.array_equal(lanes.download(), ug)


def test_grouped_set_equals_separate_contexts(mf):
    """A grouped set is the clouds side by side, nothing more: with the
    sources that act node by node (Flyer and Tominec hyperviscosity, the
    upwind viscosity, with common coefficients) each cloud's rhs! in one
    grouped context equals, bit for bit, the same cloud in a context of its
    own, in the reference's operation order and in split-FP32."""
    from mfrbf.clouds import cloud_square
    from mfrbf.ensemble import Ensemble

    specs = configs.c5(3)
    for r, s in enumerate(specs):
        s.side, s.hv = 24 + 6 * r, None
    ens = Ensemble(specs)
    sp = specs[0]
    gam_f, gam_t, h = 1e-4 * min(ens.dx_min) ** 4, 1e-3 * min(ens.dx_min) ** 4.5, float(np.mean(ens.dx_avg))

    def sources(ctx):
        ctx.compute_operator(sp.p, sp.deg, 4, L.MF_OP_HYPERVISCOSITY, want=False)
        ctx.set_source_hyperviscosity(gam_f)
        ctx.compute_operator(sp.p, sp.deg, 2, L.MF_OP_LAPLACIAN, want=False)
        ctx.set_source_hyperviscosity_tominec(gam_t)
        ctx.set_source_viscosity(L.MF_VISC_UPWIND, 1.0, 1.0, h, 3, True)

    sources(ens.ctx)
    rng = np.random.default_rng(5)
    u = ens.u0 * (1.0 + 0.01 * rng.standard_normal(ens.u0.shape))
    for prec in (L.MF_PREC_F64_REF, L.MF_PREC_SPLIT32):
        ug, dug = u.copy(), np.zeros_like(u)
        ens.ctx.rhs_host(0.3, ug, dug, prec)
        for r, spec in enumerate(specs):
            lo, hi = ens.offsets[r], ens.offsets[r + 1]
            x, y = ens.x[lo:hi], ens.y[lo:hi]
            ctx = L.Context(0)
            ctx.set_equations(L.MF_EQ_EULER_2D, spec.gamma)
            ctx.set_points(x, y)
            ctx.knn(spec.k)
            ctx.compute_operator(spec.p, spec.deg, 1, L.MF_OP_FLUX, want=False)
            tag = cloud_square(spec.side, spec.seed)[2]
            inflow, outflow = np.nonzero(tag == 1)[0].astype(np.int32), np.nonzero(tag == 2)[0].astype(np.int32)
            prm = spec.ic.params()
            ctx.set_boundary([0, len(inflow), len(inflow) + len(outflow)], [L.MF_BC_DIRICHLET, L.MF_BC_DO_NOTHING],
                             [spec.ic.code, spec.ic.code], np.concatenate([prm, prm]),
                             np.concatenate([inflow, outflow]), np.zeros((len(inflow) + len(outflow), 2)))
            sources(ctx)
            us, dus = np.ascontiguousarray(u[:, lo:hi]), np.zeros((4, hi - lo))
            ctx.rhs_host(0.3, us, dus, prec)
            _say("  prec %d cloud %d (N=%d): max |du_grouped - du_alone| %.1e" % (
                prec, r, hi - lo, np.abs(dug[:, lo:hi] - dus).max()))
            assert np.array_equal(dug[:, lo:hi], dus)
            assert np.array_equal(ug[:, lo:hi], us)
