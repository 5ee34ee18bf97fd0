"""GPU workflows through the host mirror of the reference API and the C ABI,
against the CPU oracle: Medusa fixture with slip walls, the Zhang-Shu
limiter (standalone and as the stage limiter), full configuration runs,
graph replay against eager stepping, and the adjoint against finite
differences."""
import os

import numpy as np
import pytest

import cases
import oracle as O
from conftest import GOLDEN, rel_inf
from mfrbf import _lib as L
from mfrbf import configs

import mfrbf

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


CYL = os.path.join(GOLDEN, "cyl_0_05", "cyl_0_05")
CYL_TAGS = {"inlet": 1, "outlet": 2, "bottom": 3, "top": 4, "cylinder": 5}


def _cyl_semi(visc=None, host_dirichlet=False):
    basis = mfrbf.PointCloudBasis(mfrbf.Point2D(), 3, approximation_type=mfrbf.RBF(mfrbf.PolyharmonicSpline(5)))
    solver = mfrbf.PointCloudSolver(basis)
    domain = mfrbf.PointCloudDomain(solver, CYL, CYL_TAGS)
    eqs = mfrbf.CompressibleEulerEquations2D(1.4)
    grad = mfrbf.LinearGradient()
    g = (lambda x, t, e: grad(x, t, e)) if host_dirichlet else grad
    bcs = {"inlet": mfrbf.BoundaryConditionDirichlet(g), "outlet": mfrbf.BoundaryConditionDoNothing(),
           "bottom": mfrbf.boundary_condition_slip_wall, "top": mfrbf.boundary_condition_slip_wall,
           "cylinder": mfrbf.boundary_condition_slip_wall}
    src = None
    if visc == "upwind":
        src = mfrbf.SourceTerms(uv=mfrbf.SourceUpwindViscosityTominec(solver, eqs, domain, c=1.0, c_uw=1.0))
    elif visc == "tominec_upwind":  # test/upwind_viscosity_test.jl:50-52
        src = mfrbf.SourceTerms(hv=mfrbf.SourceHyperviscosityTominec(solver, eqs, domain, c=1.0),
                                rv=mfrbf.SourceUpwindViscosityTominec(solver, eqs, domain))
    elif visc == "igr":
        src = mfrbf.SourceTerms(igr=mfrbf.SourceIGR(solver, eqs, domain, alpha=domain.pd.dx_min ** 2))
    elif visc == "flyer_tominec":  # test/hyperviscosity_test.jl:42-45
        src = mfrbf.SourceTerms(hv=mfrbf.SourceHyperviscosityFlyer(solver, eqs, domain, k=2, c=1.0),
                                hv2=mfrbf.SourceHyperviscosityTominec(solver, eqs, domain, c=1.0))
    semi = mfrbf.SemidiscretizationHyperbolic(domain, eqs, grad, solver, boundary_conditions=bcs,
                                              source_terms=src)
    return semi, domain


def _cyl_oracle(domain, visc=None):
    pd = domain.pd
    x, y = pd.points[:, 0].copy(), pd.points[:, 1].copy()
    N, k = pd.num_points, pd.num_neighbors
    idx, _ = O.knn_brute(x, y, k)
    assert np.array_equal(idx, pd.neighbors)  # GPU kNN of the host API is exact
    wx, wy, rc = O.weights(x, y, O.ell_ptr(N, k), idx, 5, 3, 1)
    assert rc == 0
    tags = domain.boundary_tags
    boundary = [(O.BC_DIRICHLET, tags["inlet"].idx, tags["inlet"].normals, O.FN_GRADIENT, None, None),
                (O.BC_NOTHING, tags["outlet"].idx, tags["outlet"].normals, None, None, None)]
    for name in ("bottom", "top", "cylinder"):
        boundary.append((O.BC_SLIP, tags[name].idx, tags[name].normals, None, None, None))
    rv = O.RV_UPWIND if visc in ("upwind", "tominec_upwind") else O.RV_NONE
    kw = {}
    if visc in ("tominec_upwind", "flyer_tominec"):
        lx, ly, rc = O.weights(x, y, O.ell_ptr(N, k), idx, 5, 3, 2)
        kw.update(lap=(lx + ly).reshape(N, k), tom_gamma=pd.dx_min ** 4.5)
    if visc == "tominec_upwind":
        kw.update(order=[O.SRC_TOMINEC, O.SRC_VISCOSITY])
    if visc == "igr":
        kw.update(igr_alpha=pd.dx_min ** 2, order=[O.SRC_IGR])
    if visc == "flyer_tominec":
        hx, hy, rc = O.weights(x, y, O.ell_ptr(N, k), idx, 5, 3, 4)
        kw.update(hv=(hx + hy).reshape(N, k), hv_gamma=pd.dx_min ** 4, order=[O.SRC_FLYER, O.SRC_TOMINEC])
    return O.Problem(x, y, O.ell_ptr(N, k), idx, wx, wy, O.EQ_EULER, 4, boundary=boundary, rv_mode=rv,
                     c_rv=1.0, c_uw=1.0, dx_avg=pd.dx_avg, **kw)


@pytest.mark.parametrize("visc,host_dirichlet", [(None, False), ("upwind", False), (None, True),
                                                 ("tominec_upwind", False), ("flyer_tominec", False),
                                                 ("igr", False)])
def test_host_api_cyl_slip_walls(mf, visc, host_dirichlet):
    """Trixi.rhs! on the reference's Medusa cloud (2,154 nodes, five boundary
    groups): Dirichlet inlet, do-nothing outlet, three slip walls; with the
    source combinations of the reference's viscosity tests."""
    _say("host API cyl_0_05 visc=%s host_dirichlet=%s" % (visc, host_dirichlet))
    semi, domain = _cyl_semi(visc, host_dirichlet)
    prob = _cyl_oracle(domain, visc)
    ode = mfrbf.semidiscretize(semi, (0.0, 0.1))
    u = ode.u0.copy()
    rng = np.random.default_rng(7)
    u[1:3] += 0.05 * rng.standard_normal(u[1:3].shape)  # momentum not tangent at the walls
    uo = u.copy()
    du = np.zeros_like(u)
    mfrbf.rhs_(du, u, semi, 0.02)
    duo, _, _ = prob.rhs(uo, 0.02)
    assert rel_inf(u, uo) < 1e-15  # boundary passes mutate u identically
    assert rel_inf(du, duo) < 1e-12
    for name in ("bottom", "top", "cylinder"):
        b = domain.boundary_tags[name]
        nrm = b.normals / np.linalg.norm(b.normals, axis=1)[:, None]
        assert np.abs((u[1:3, b.idx].T * nrm).sum(1)).max() < 1e-14
        assert np.all(du[1:3, b.idx] == 0.0)
    assert np.all(du[:, domain.boundary_tags["inlet"].idx] == 0.0)


def test_trixi2vtk_cyl(mf, tmp_path):
    """trixi2vtk(u, semi, t) after a rhs! with the upwind viscosity: the file
    holds the state, the primitive variables, eps and its code, the 1-based
    index and the time; a second call with iter = 1 appends to the .pvd."""
    _say("trixi2vtk on cyl_0_05")
    from mfrbf import vtk
    semi, domain = _cyl_semi("upwind")
    ode = mfrbf.semidiscretize(semi, (0.0, 0.1))
    u = ode.u0.copy()
    du = np.zeros_like(u)
    mfrbf.rhs_(du, u, semi, 0.0)
    path = mfrbf.trixi2vtk(u, semi, 0.0, iter=0, output_directory=str(tmp_path), prefix="sol")
    assert os.path.basename(path) == "sol_CompressibleEulerEquations2D_1_0.vtu"
    p, pd, fd = vtk.read_vtu(path)
    assert np.array_equal(p[:, :2], domain.pd.points)
    assert np.array_equal(pd["density"], u[0]) and np.array_equal(pd["density_energy"], u[3])
    assert np.array_equal(pd["momentum"], u[1:3].T)
    assert np.allclose(pd["velocity"], (u[1:3] / u[0]).T, rtol=1e-15)
    eps, code = semi.ctx.viscosity()
    assert np.array_equal(pd["eps"], eps) and np.array_equal(pd["eps_scalar"], code)
    assert np.array_equal(pd["index"], np.arange(1, u.shape[1] + 1))
    assert float(fd["time"][0]) == 0.0
    mfrbf.trixi2vtk(u, semi, 0.5, iter=1, output_directory=str(tmp_path), prefix="sol")
    pvd = open(os.path.join(str(tmp_path), "sol_CompressibleEulerEquations2D_1.pvd")).read()
    assert pvd.count("<DataSet") == 2 and 'timestep="0.5"' in pvd
    # the IGR source writes sigma
    semi2, _ = _cyl_semi("igr")
    mfrbf.rhs_(du, u.copy(), semi2, 0.0)
    _, pd2, _ = vtk.read_vtu(mfrbf.trixi2vtk(u, semi2, 0.0, output_directory=str(tmp_path), prefix="igr"))
    assert np.array_equal(pd2["sigma"], semi2.ctx.igr_sigma()) and np.abs(pd2["sigma"]).max() > 0.0


def test_limiter_vs_oracle(mf):
    """Trixi.limiter_zhang_shu! (positivity_zhang_shu_point2d.jl:22-83) on
    density and then pressure, with nodes pushed below the thresholds."""
    _say("zhang-shu limiter")
    c = cases.build(configs.c2(), side=40)
    u = c.u0.copy()
    rng = np.random.default_rng(3)
    bad = rng.choice(c.N, 60, replace=False)
    u[0, bad[:30]] = 0.05
    u[3, bad[30:]] = 0.2 * u[3, bad[30:]]  # low pressure
    ptr = O.ell_ptr(c.N, c.k)
    for thr, var in ((0.5, L.MF_VAR_DENSITY), (0.5, L.MF_VAR_PRESSURE)):
        uo = O.limiter_zhang_shu(u.copy(), ptr, c.idx, thr, var)
        changed = np.any(uo != u, axis=0)
        assert changed.sum() > 0
        out = []
        for tiles in (0, 1):  # global gathers; the stencil kernel's LDS tiles (built by a rhs! call)
            c.ctx.set_tuning(L.MF_TUNE_TILES, tiles)
            c.ctx.rhs_host(0.0, c.u0.copy(), np.zeros_like(c.u0))
            assert c.ctx.tile_info()["tiles_a"] == tiles
            c.ctx.upload(L.MF_BUF_U, u.copy())
            c.ctx.limiter_zhang_shu(thr, var)
            out.append(c.ctx.download(L.MF_BUF_U))
            assert rel_inf(out[-1], uo) < 1e-14
        assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("k", [15, 13])
def test_limiter_scalar_vs_oracle(mf, k):
    """The limiter on a scalar state (advection, C1 stencils; the variable
    is the state itself), with a stencil size that is not a multiple of the
    10-entry load batch."""
    _say("zhang-shu limiter scalar k=%d" % k)
    c = cases.build(configs.c1(), side=40)
    if k != c.k:
        idx, _ = O.knn_grid(c.x, c.y, k)
        c.ctx.set_stencil(idx)
        c.idx, c.k = idx, k
    u = c.u0.copy()
    rng = np.random.default_rng(4)
    bad = rng.choice(c.N, 40, replace=False)
    u[0, bad] = -0.2
    ptr = O.ell_ptr(c.N, c.k)
    ug = u.copy()
    c.ctx.upload(L.MF_BUF_U, ug)
    c.ctx.limiter_zhang_shu(0.0, L.MF_VAR_DENSITY)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = O.limiter_zhang_shu(u.copy(), ptr, c.idx, 0.0, L.MF_VAR_DENSITY)
    assert np.any(uo != u)
    assert rel_inf(ug, uo) < 1e-14


def test_ssprk33_stage_limiter_vs_oracle(mf):
    _say("ssprk33 with stage limiter")
    c = cases.build(configs.c2(), side=40)
    u0 = c.u0.copy()
    u0[0, ::97] = 0.3
    lim = [(0.6, O.VAR_DENSITY), (0.5, O.VAR_PRESSURE)]
    c.ctx.set_stage_limiter([t for t, _ in lim], [L.MF_VAR_DENSITY, L.MF_VAR_PRESSURE])
    c.ctx.upload(L.MF_BUF_U, u0.copy())
    c.ctx.ssprk33(0.0, c.dt, 5, flags=L.MF_STEP_LIMITER)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 5, limiter=lim, stencil=(O.ell_ptr(c.N, c.k), c.idx))
    assert rel_inf(ug, uo) < 1e-12


def test_ssprk33_c1_full_run(mf):
    """C1: advection of a Gaussian pulse, 4,096 nodes, 200 SSPRK33 steps."""
    _say("C1 200 steps")
    spec = configs.c1()
    c = cases.build(spec)
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.ssprk33(0.0, c.dt, spec.nsteps, flags=L.MF_STEP_GRAPH)
    ug = c.ctx.download(L.MF_BUF_U)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, spec.nsteps)
    err = rel_inf(ug, uo)
    _say("  C1 rel err %.2e" % err)
    assert err < 1e-11


@pytest.mark.parametrize("prec", [L.MF_PREC_F64, L.MF_PREC_SPLIT32])
def test_graph_replay_matches_eager(mf, prec):
    """The captured step, replayed over several calls that continue the
    trajectory, gives the same bits as eager stepping."""
    _say("graph replay vs eager prec %d" % prec)
    c = cases.build(configs.c2(), side=48, visc="residual", hv=(2, 1e-4))
    flags = L.MF_STEP_HISTORY
    out = {}
    for graph in (False, True):
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        t = 0.0
        for call in range(4):
            c.ctx.ssprk33(t, c.dt, 3, prec, flags | (L.MF_STEP_GRAPH if graph else 0), resume=(call > 0))
            t += 3 * c.dt
            # eager work between replays
            c.ctx.synchronize()
        out[graph] = c.ctx.download(L.MF_BUF_U)
    assert np.array_equal(out[False], out[True])


def test_multi_step_graph_and_new_dt(mf):
    """20 steps as one 16-step graph plus four 1-step graphs give the eager
    bits; a later call with another dt recaptures instead of replaying the
    old step."""
    _say("multi-step graph, then a new dt")
    c = cases.build(configs.c2(), side=40, hv=(2, 1e-4))
    out = {}
    for graph in (False, True):
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.ssprk33(0.0, c.dt, 20, L.MF_PREC_F64, L.MF_STEP_GRAPH if graph else 0)
        a = c.ctx.download(L.MF_BUF_U)
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.ssprk33(0.0, 0.5 * c.dt, 17, L.MF_PREC_F64, L.MF_STEP_GRAPH if graph else 0)
        out[graph] = (a, c.ctx.download(L.MF_BUF_U))
    assert np.array_equal(out[False][0], out[True][0])
    assert np.array_equal(out[False][1], out[True][1])
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, 0.5 * c.dt, 17)
    assert rel_inf(out[True][1], uo) < 1e-12


def test_adjoint_vs_finite_differences_c4(mf):
    """C4: dL/drho0 of L = sum rho(T)^2 through 100 SSPRK33 steps on 65,536
    nodes (checkpoints every 10) against central differences with the
    configuration's step eps = 1e-6 along random directions and single
    coordinates, within 1e-6 relative.  The random directions have N(0, 1)
    entries (each node's density moves by about eps), not unit length: with
    a unit vector spread over 65,536 nodes each node moves by 4e-9 and the
    rounding of the 300 stages (about 1e-14 per node) is 1e-6 of the
    difference.  L(u+) - L(u-) is summed as
    sum (rho+ - rho-)(rho+ + rho-), the same number without the cancellation
    of two sums of 65,536 squares (whose rounding, about 1e-11, would
    swamp a difference of order 1e-6)."""
    _say("C4 adjoint vs finite differences")
    from bench import setup_euler
    spec = configs.c4()
    ctx, info = setup_euler(L, spec, spec.side, visc=False)
    N, dt = info["N"], info["dt"]
    u0 = configs.perturbed_density(info["u0"], spec, N)
    ctx.upload(L.MF_BUF_U, u0)
    loss = ctx.ssprk33_vjp(0.0, dt, spec.nsteps, spec.checkpoint_every)
    grad = ctx.download(L.MF_BUF_AUX1)[0]

    def rho_T(u):
        ctx.upload(L.MF_BUF_U, u)
        ctx.ssprk33(0.0, dt, spec.nsteps)
        return ctx.download(L.MF_BUF_U)[0]

    def L_of(u):
        return float(np.sum(rho_T(u) ** 2))

    assert abs(L_of(u0) - loss) <= 1e-13 * abs(loss)
    rng = np.random.default_rng(spec.noise_seed)
    h = 1e-6
    errs = []
    dirs = [rng.standard_normal(N) for _ in range(8)]
    for i in rng.choice(N, 8, replace=False):
        e = np.zeros(N)
        e[i] = 1.0
        dirs.append(e)
    for d in dirs:
        up, um = u0.copy(), u0.copy()
        up[0] += h * d
        um[0] -= h * d
        rp, rm = rho_T(up), rho_T(um)
        fd = float(np.sum((rp - rm) * (rp + rm))) / (2 * h)
        ad = float(grad @ d)
        errs.append(abs(fd - ad) / max(abs(ad), 1e-300))
    _say("  FD rel errors max %.2e" % max(errs))
    assert max(errs) < 1e-6


def test_ensemble_nosync_matches_one_at_a_time(mf):
    """Contexts stepped concurrently (MF_STEP_NOSYNC, one stream each) give
    the same bits as the same steps one context at a time."""
    _say("ensemble: concurrent vs one at a time")
    specs = configs.c5(3)
    cs = []
    for spec in specs:
        spec.side = min(spec.side, 96)
        c = cases.build(spec, hv=(2, 1e-4), bc=("dirichlet", "nothing"))
        cs.append(c)
    res = {}
    for conc in (True, False):
        for c in cs:
            c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        for c in cs:
            f = L.MF_STEP_GRAPH | (L.MF_STEP_NOSYNC if conc else 0)
            c.ctx.ssprk33(0.0, c.dt, 7, L.MF_PREC_SPLIT32, f)
        for c in cs:
            c.ctx.synchronize()
        res[conc] = [c.ctx.download(L.MF_BUF_U) for c in cs]
    for a, b in zip(res[True], res[False]):
        assert np.array_equal(a, b)


def test_c_client(mf, tmp_path):
    """A plain C99 host (tests/c_client/mfrbf_client.c) drives kNN, weights,
    rhs! and graph-replayed SSPRK33 steps through the C ABI and checks each
    against a closed form (linear exactness, quadratic reproduction, free
    stream); no Python in its loop."""
    import subprocess

    from conftest import build_c_client

    exe = build_c_client(tmp_path)
    _say("C client")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    _say("  " + r.stdout.strip() + r.stderr.strip())
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "mfrbf C client ok" in r.stdout


def test_small_cloud_igr_limiter_adjoint(mf):
    """A cloud smaller than one 256-node tile (12 x 12 = 144 nodes, 30-point
    stencils): the IGR source (alpha = dx_min^2), the Zhang-Shu limiter on
    density and pressure, and the VJP of rhs!, each against the oracle."""
    _say("small cloud: IGR, limiter, VJP")
    spec = configs.c2()
    c0 = cases.build(spec, side=12, use_gpu=False)
    c = cases.build(spec, side=12, igr=c0.dx_min ** 2)
    u = c.u0.copy()
    du = np.zeros_like(u)
    c.ctx.rhs_host(0.01, u, du, L.MF_PREC_F64_REF)
    duo, _, _ = c.prob.rhs(c.u0.copy(), 0.01)
    assert rel_inf(c.ctx.igr_sigma(), c.prob.igr_sigma) < 1e-10
    assert rel_inf(du, duo) < 1e-10
    ub = c.u0.copy()
    rng = np.random.default_rng(4)
    bad = rng.choice(c.N, 12, replace=False)
    ub[0, bad[:6]] = 0.05
    ub[3, bad[6:]] = 0.2 * ub[3, bad[6:]]
    ptr = O.ell_ptr(c.N, c.k)
    for thr, var in ((0.5, L.MF_VAR_DENSITY), (0.5, L.MF_VAR_PRESSURE)):
        uo = O.limiter_zhang_shu(ub.copy(), ptr, c.idx, thr, var)
        assert np.any(uo != ub)
        c.ctx.upload(L.MF_BUF_U, ub.copy())
        c.ctx.limiter_zhang_shu(thr, var)
        assert rel_inf(c.ctx.download(L.MF_BUF_U), uo) < 1e-14
    cv = cases.build(spec, side=12)
    lam = rng.standard_normal(cv.u0.shape)
    cv.ctx.upload(L.MF_BUF_U, cv.u0.copy())
    cv.ctx.upload(L.MF_BUF_AUX0, lam)
    cv.ctx.rhs_vjp(0.0)
    assert rel_inf(cv.ctx.download(L.MF_BUF_AUX1), cv.prob.vjp(cv.u0.copy(), lam, 0.0)) < 1e-12


@pytest.mark.parametrize("order", [1, 2, 5, 7])
def test_ssprk33_rv_history_orders(mf, order):
    """HistoryCallback(approx_order) for orders other than the
    configurations' 3, up to the largest history kept (8 states): 10
    graph-replayed SSPRK33 steps with the residual viscosity, reference
    order, against the oracle (state 1e-12; approx_du, a time difference
    with weights of order 1/dt, 1e-8)."""
    _say("ssprk33 + residual viscosity, approx_order=%d" % order)
    spec = configs.c2()
    spec.approx_order = order
    c = cases.build(spec, side=40, visc="residual")
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.ssprk33(0.0, c.dt, 10, L.MF_PREC_F64_REF, L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 10, approx_order=order)
    e_u = rel_inf(c.ctx.download(L.MF_BUF_U), uo)
    e_ad = rel_inf(c.ctx.approx_du().ravel(), c.prob.approx_du)
    _say("  state %.2e, approx_du %.2e" % (e_u, e_ad))
    assert np.abs(c.prob.approx_du).max() > 0.0
    assert e_u < 1e-12 and e_ad < 1e-8


def test_rhs_igr_zero_iterations(mf):
    """The IGR source with no CG iteration: sigma keeps its starting value
    and du its flux terms, as the oracle's."""
    _say("IGR with maxiter = 0")
    spec = configs.c2()
    c0 = cases.build(spec, side=32, use_gpu=False)
    c = cases.build(spec, side=32, igr=c0.dx_min ** 2, igr_maxiter=0)
    u = c.u0.copy()
    du = np.zeros_like(u)
    c.ctx.rhs_host(0.01, u, du, L.MF_PREC_F64_REF)
    duo, _, _ = c.prob.rhs(c.u0.copy(), 0.01)
    assert np.array_equal(c.ctx.igr_sigma(), c.prob.igr_sigma)
    assert rel_inf(du, duo) < 1e-12


def test_limiter_every_node_below_threshold(mf):
    """A density threshold above every node's density: every row is blended
    with its stencil mean, as the oracle does."""
    _say("limiter with every node below the threshold")
    c = cases.build(configs.c2(), side=24)
    u = c.u0.copy()
    ptr = O.ell_ptr(c.N, c.k)
    thr = 1.5 * u[0].max()
    uo = O.limiter_zhang_shu(u.copy(), ptr, c.idx, thr, O.VAR_DENSITY)
    assert np.all(np.any(uo != u, axis=0))
    c.ctx.upload(L.MF_BUF_U, u.copy())
    c.ctx.limiter_zhang_shu(thr, L.MF_VAR_DENSITY)
    assert rel_inf(c.ctx.download(L.MF_BUF_U), uo) < 1e-14


def test_boundary_node_in_two_tags(mf):
    """Nodes in two boundary tags (the bottom edge: in the Dirichlet inflow
    tag, then in a slip-wall tag) take both, in tag order, as the
    reference's per-tag loop does (pointcloudsolver/rbfsolver.jl:277-286):
    rhs! in the three precisions, and 4 graph-replayed SSPRK33 steps with
    the residual viscosity and the history callback (each rhs! then runs its
    own boundary and sum passes), against the oracle."""
    _say("boundary nodes in two tags")
    spec = configs.c2()
    c = cases.build(spec, side=32, visc="residual")
    inflow = np.nonzero(c.tag == 1)[0].astype(np.int32)
    outflow = np.nonzero(c.tag == 2)[0].astype(np.int32)
    bottom = inflow[c.y[inflow] == 0.0]
    assert len(bottom) > 3
    nrm = np.tile([0.0, -1.0], (len(bottom), 1))
    fn, prm = spec.ic.code, spec.ic.params()
    ptr = [0, len(inflow), len(inflow) + len(bottom), len(inflow) + len(bottom) + len(outflow)]
    c.ctx.set_boundary(ptr, [L.MF_BC_DIRICHLET, L.MF_BC_SLIP_WALL, L.MF_BC_DO_NOTHING], [fn, 0, 0],
                       np.concatenate([prm, prm, prm]), np.concatenate([inflow, bottom, outflow]),
                       np.concatenate([np.zeros((len(inflow), 2)), nrm, np.zeros((len(outflow), 2))]))
    prob = O.Problem(c.x, c.y, c.ptr, c.idx, c.wx.ravel(), c.wy.ravel(), c.eq, c.nvar, gamma=spec.gamma,
                     boundary=[(O.BC_DIRICHLET, inflow, None, fn, prm, None), (O.BC_SLIP, bottom, nrm, fn, prm, None),
                               (O.BC_NOTHING, outflow, None, fn, prm, None)],
                     rv_mode=O.RV_RESIDUAL, dx_avg=c.dx_avg)
    u0 = c.u0.copy()
    u0[1:3] += 0.02 * np.random.default_rng(12).standard_normal(u0[1:3].shape)
    duo, _, _ = prob.rhs(u0.copy(), 0.02)
    for prec, tol in ((L.MF_PREC_F64, 1e-12), (L.MF_PREC_F64_REF, 1e-12), (L.MF_PREC_SPLIT32, 1e-6)):
        u = u0.copy()
        du = np.zeros_like(u)
        c.ctx.rhs_host(0.02, u, du, prec)
        assert rel_inf(du, duo) < tol, prec
    c.ctx.upload(L.MF_BUF_U, u0.copy())
    c.ctx.ssprk33(0.0, c.dt, 4, L.MF_PREC_F64_REF, L.MF_STEP_HISTORY | L.MF_STEP_GRAPH)
    uo = u0.copy()
    prob.ssprk33(uo, 0.0, c.dt, 4, approx_order=3)
    assert rel_inf(c.ctx.download(L.MF_BUF_U), uo) < 1e-12


@pytest.mark.parametrize("bc", [("nothing", "nothing"), ("dirichlet", "slip")])
def test_adjoint_checkpoint_intervals_agree(mf, bc):
    """The reverse sweep recomputes each segment from its checkpoint with
    the forward's own kernels, so the gradient does not depend on the
    checkpoint interval: 7 SSPRK33 steps with checkpoints every 7 (one
    segment, the tape of the forward pass), 3 (a partial last segment), 2
    and 1 step give the same loss and dL/du0 bit for bit."""
    _say("adjoint checkpoint intervals, bc=%s" % (bc,))
    c = cases.build(configs.c2(), side=24, bc=bc)
    u0 = c.u0.copy()
    u0[1:3] += 0.02 * np.random.default_rng(8).standard_normal(u0[1:3].shape)
    outs = []
    for every in (7, 3, 2, 1):
        c.ctx.upload(L.MF_BUF_U, u0.copy())
        loss = c.ctx.ssprk33_vjp(0.0, c.dt, 7, every)
        outs.append((loss, c.ctx.download(L.MF_BUF_AUX1)))
    assert np.abs(outs[0][1]).max() > 0.0
    for loss, g in outs[1:]:
        assert loss == outs[0][0] and np.array_equal(g, outs[0][1])
