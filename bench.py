#!/usr/bin/env python3
"""RHS throughput of the RBF-FD semidiscretization on one MI355X.

Workload (the headline configuration): compressible Euler (gamma = 1.4) with
residual viscosity on 262,144 jittered nodes (512 x 512, +-0.3h, seeded),
PHS r^5 + degree-3 polynomials, 30-point stencils, density-wave state,
Dirichlet inflow / do-nothing outflow.  One step = one rhs! over the whole
cloud (262,144 node-evaluations) on device-resident state, FP64 in the
reference's operation order (the mode within 1e-12 of the reference).

Printed as one JSON line, with:
  roofline      the dominant kernel's algorithmic bytes / its event-timed
                duration vs 8 TB/s (whole-rhs! fraction reported beside it)
  cpu_baseline  the FP64 CPU oracle timed on this host with OpenMP (one
                thread per core of the host's share), 1-thread rate beside it
  f64_fma       the fused-multiply-add FP64 rhs! (faster, within 5e-12)
  split32       the split-FP32 rhs! rate, speed-up and error vs FP64
  adjoint       reverse-sweep / forward time of 100 SSPRK33 steps (65,536 nodes)
  c1_advection  200 SSPRK33 steps of advection on 4,096 nodes, GPU and CPU oracle
  cold_cache    the same rhs! after a flush of the L2 and last-level caches
  weights       batched weight generation for 1,048,576 ragged stencils,
                with the CPU oracle on a 32,768-stencil sample beside it
  ensemble      64 clouds (5.7M nodes), 50 split-FP32 SSPRK33 steps each
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "meshfreetrixi.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "RHS node-evals/sec: FP64 vs split-FP32 at ≤1e-6 rel err; adjoint/fwd time"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def setup_euler(L, spec, side, visc=True):
    from mfrbf.clouds import cloud_square
    from mfrbf.configs import euler_dt
    from mfrbf.equations import CompressibleEulerEquations2D

    x, y, tag, nx, ny = cloud_square(side, spec.seed)
    N = len(x)
    ctx = L.Context(0)
    ctx.set_equations(L.MF_EQ_EULER_2D, spec.gamma)
    ctx.set_points(x, y)
    t0 = time.perf_counter()
    idx, _, dx_min, dx_avg = ctx.knn(spec.k)
    t_knn = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.compute_operator(spec.p, spec.deg, 1, L.MF_OP_FLUX, want=False)
    t_w = time.perf_counter() - t0
    inflow = np.nonzero(tag == 1)[0].astype(np.int32)
    outflow = np.nonzero(tag == 2)[0].astype(np.int32)
    prm = spec.ic.params()
    ctx.set_boundary([0, len(inflow), len(inflow) + len(outflow)], [L.MF_BC_DIRICHLET, L.MF_BC_DO_NOTHING],
                     [spec.ic.code, spec.ic.code], np.concatenate([prm, prm]),
                     np.concatenate([inflow, outflow]), np.zeros((N, 2))[:len(inflow) + len(outflow)])
    if visc:
        ctx.set_source_viscosity(L.MF_VISC_RESIDUAL, 1.0, 1.0, dx_avg, 3, True)
    eq = CompressibleEulerEquations2D(spec.gamma)
    u0 = np.ascontiguousarray(spec.ic(np.c_[x, y], 0.0, eq))
    dt = euler_dt(spec, u0, dx_min)
    info = dict(x=x, y=y, tag=tag, idx=idx, dx_min=dx_min, dx_avg=dx_avg, u0=u0, dt=dt, N=N,
                t_knn=t_knn, t_weights=t_w, inflow=inflow, outflow=outflow)
    return ctx, info


def rhs_bytes(N, k, V=4, visc=True):
    """Algorithmic bytes of one rhs!: each array counted once, int32 indices,
    FP64 weights; the intermediate G is not counted."""
    flux = N * (V * 8 + V * 8 + k * 4 + k * 16)
    if visc:
        flux += N * (k * 4 + k * 16 + 4) + N * V * 8  # transposed operator + approx_du
    return flux


def kernel_bytes(N, k, ti, V=4):
    """Bytes the stencil kernel (phase A) and the transposed gather (phase B)
    must move in the format they use, each array counted once (gathers of
    the halo not counted; the reference-order kernel reads its stencil
    positions once for the x pass and again for the y pass, but the array is
    counted once).  Phase A reads u, the stencil positions, (Dx, Dy) and
    approx_du and writes du, G, eps and its code; phase B reads the
    transposed positions and (Dx, Dy) (N k entries, padding not counted), G
    and du and writes du.  With LDS tiles the positions are 16-bit and each
    tile's halo list (int32) and offsets are read once; without, the
    positions are int32 node indices."""
    if ti["tiles_a"]:
        a = N * (V * 8 + 2 * k + 16 * k + V * 8 + V * 8 + 2 * V * 8 + 9) + 4 * ti["halo_a_total"] \
            + 8 * (ti["n_tiles"] + 1)
    else:
        a = N * (V * 8 + 4 * k + 16 * k + V * 8 + V * 8 + 2 * V * 8 + 9)
    if ti["tiles_b"]:
        b = N * k * (2 + 16) + N * (2 * V * 8 + V * 8 + V * 8 + 1) + 4 * ti["halo_b_total"] \
            + ti["n_tiles"] * (256 + 4 * (8 + 4) + 8)
    else:
        b = N * (4 * k + 16 * k + 2 * V * 8 + V * 8 + V * 8 + 1)
    return a, b


def measured_traffic(kernel, tuning, side, tiled, mode="f64_ref"):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (FETCH_SIZE
    scaled by the copy-kernel calibration recorded beside it), or None."""
    path = os.path.join(ROOT, "profiles", "rhs_euler_rv_262144_hbm_traffic.json")
    if side != 512 or not os.path.exists(path):
        return None, None
    with open(path) as f:
        tab = json.load(f)
    tl = "true" if tiled else "false"
    tm = 2 if mode == "f64_ref" else 0
    if kernel == "transpose":
        want = "k_rhs_transpose<4, %d, %d, %s, 1>" % (tm, tuning[mode]["transpose"] or 4, tl)
        family, tail = "k_rhs_transpose<4, %d," % tm, "%s, 1>(mf::RhsArgs)" % tl
    else:
        kern = "k_rhs_ref" if mode == "f64_ref" else "k_rhs_f64"
        want = "%s<4, 30, 0, true, %d, %s>" % (kern, tuning[mode]["flux"] or 5, tl)
        family, tail = "%s<4, 30, 0, true," % kern, "%s>(mf::RhsArgs)" % tl
    names = [n for n in tab["kernels"] if want in n] or \
        [n for n in tab["kernels"] if family in n and n.endswith(tail)]
    if not names:
        return None, None
    return tab["kernels"][names[0]]["bytes"], names[0]


def tominec(L, ctx, info, spec, t_now, steps):
    """The same rhs! with the Tominec hyperviscosity added after the residual
    viscosity (lap computed on the device, gamma = dx_min^4.5): per-precision
    rhs! time, kernel times and the two kernels' HBM fractions.  The Tominec
    source adds to phase A its lap weights (N k doubles) and the store of
    L = lap u (N V doubles), to phase B the lap of the transposed entries
    (N k doubles) and the read of L (N V doubles)."""
    N, k, V = info["N"], spec.k, 4
    ctx.compute_operator(spec.p, spec.deg, 2, L.MF_OP_LAPLACIAN, want=False)
    ctx.set_source_hyperviscosity_tominec(info["dx_min"] ** 4.5)
    ctx.set_source_order([L.MF_SRC_VISCOSITY, L.MF_SRC_TOMINEC])
    res = {}
    try:
        for name, prec in (("f64_ref", L.MF_PREC_F64_REF), ("f64_fma", L.MF_PREC_F64),
                           ("split32", L.MF_PREC_SPLIT32)):
            ctx.time_rhs(t_now, prec, 20)
            ms, parts = ctx.time_rhs(t_now, prec, steps)
            ti = ctx.tile_info()
            ba, bb = kernel_bytes(N, k, ti)
            ba += N * k * 8 + N * V * 8
            bb += N * k * 8 + N * V * 8
            res[name] = {"value": N / (ms * 1e-3), "ms_per_step": ms,
                         "kernels_ms": {"stencil": parts[0], "transpose": parts[1], "reductions": parts[2]}}
            if name != "split32":
                res[name]["stencil_frac"] = ba / (parts[0] * 1e-3) / 1e9 / HBM_PEAK_GBS
                res[name]["transpose_frac"] = bb / (parts[1] * 1e-3) / 1e9 / HBM_PEAK_GBS
        res["lds_tiles"] = ctx.tile_info()
    finally:
        ctx.set_source_hyperviscosity_tominec(0.0)
    return res


def igr(L, ctx, info, spec, t_now, steps):
    """The same rhs! with the IGR source added last (alpha = dx_min^2, 20
    conjugate-gradient iterations): rhs! time per precision.  Each CG
    iteration is two stencil passes over scalar fields and two vector
    updates, four launches; the dot products are per-block partials that the
    next launch's blocks sum."""
    ctx.set_source_igr(info["dx_min"] ** 2, 20)
    ctx.set_source_order([L.MF_SRC_VISCOSITY, L.MF_SRC_IGR])
    res = {"alpha": info["dx_min"] ** 2, "cg_iterations": 20}
    try:
        for name, prec in (("f64_ref", L.MF_PREC_F64_REF), ("f64_fma", L.MF_PREC_F64),
                           ("split32", L.MF_PREC_SPLIT32)):
            ctx.time_rhs(t_now, prec, 5)
            ms, _ = ctx.time_rhs(t_now, prec, steps)
            res[name] = {"value": info["N"] / (ms * 1e-3), "ms_per_step": ms}
    finally:
        ctx.set_source_igr(0.0, 20)
    return res


def stepping(L, ctx, info, steps):
    """SSPRK33 steps at this workload (graph-replayed, FP64 reference order):
    plain, with the history callback, and with the history callback and the
    Zhang-Shu stage limiter on density and pressure (three limiter passes per
    step), so the callbacks' cost per step is the difference; and the cost of
    one stage (a third of a step)."""
    dt = info["dt"]
    out = {}
    ctx.set_stage_limiter([1e-6, 1e-6], [L.MF_VAR_DENSITY, L.MF_VAR_PRESSURE])
    for name, flags in (("plain", 0), ("history", L.MF_STEP_HISTORY),
                        ("history_limiter", L.MF_STEP_HISTORY | L.MF_STEP_LIMITER)):
        ctx.upload(L.MF_BUF_U, info["u0"])
        ctx.ssprk33(0.0, dt, 16, L.MF_PREC_F64_REF, flags | L.MF_STEP_GRAPH)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.ssprk33(0.0, dt, steps, L.MF_PREC_F64_REF, flags | L.MF_STEP_GRAPH, resume=True)
        ctx.synchronize()
        out[name + "_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / steps
    # one stage = one rhs! with the stage update fused into its last kernel
    # (plus, with the callbacks, their share of the step)
    out["per_stage_ms"] = {name: out[name + "_ms_per_step"] / 3 for name in ("plain", "history", "history_limiter")}
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(L, ctx, info, spec, t_now, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    N, k = info["N"], spec.k
    wx, wy = ctx.compute_operator(spec.p, spec.deg, 1, L.MF_OP_FLUX, want=True, k=k)
    prm = spec.ic.params()
    prob = O.Problem(info["x"], info["y"], O.ell_ptr(N, k), info["idx"], wx, wy, O.EQ_EULER, 4,
                     gamma=spec.gamma,
                     boundary=[(O.BC_DIRICHLET, info["inflow"], None, O.FN_WAVE, prm, None),
                               (O.BC_NOTHING, info["outflow"], None, None, None, None)],
                     rv_mode=O.RV_RESIDUAL, c_rv=1.0, c_uw=1.0, dx_avg=info["dx_avg"])
    prob.approx_du[:] = ctx.approx_du().ravel()
    prob.cfg.success_iter = 3
    u = ctx.download(L.MF_BUF_U)
    du, _, _ = prob.rhs(u.copy(), t_now)  # warm
    ts, t0 = [], time.perf_counter()
    while len(ts) < 20 and time.perf_counter() - t0 < budget_s:
        uu = u.copy()
        c0 = time.perf_counter()
        du, _, _ = prob.rhs(uu, t_now)
        ts.append(time.perf_counter() - c0)
    cpu = _cpu_model()
    single = dict(value=N / float(np.median(ts)), unit="node-evals/s", cores=1, kind="port",
                  sample="median of %d FP64 rhs! calls of the full workload (%d nodes) by the C oracle, "
                  "1 thread, %s" % (len(ts), N, cpu))
    # the same rhs! with OpenMP over the host's share of cores (OMP_NUM_THREADS
    # where set): fluxes, row gathers and column dots split over threads
    nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    dp = prob.rhs_par(u.copy(), t_now, nth)  # warm: threads started, pages touched
    ts, t0 = [], time.perf_counter()
    while len(ts) < 50 and time.perf_counter() - t0 < budget_s:
        uu = u.copy()
        c0 = time.perf_counter()
        dp = prob.rhs_par(uu, t_now, nth)
        ts.append(time.perf_counter() - c0)
    par_err = float(np.abs(dp - du).max() / np.abs(du).max())
    if par_err > 1e-10:
        raise RuntimeError("OpenMP oracle rhs! differs from the serial one: %.3e" % par_err)
    return dict(value=N / float(np.median(ts)), unit="node-evals/s", cores=nth, kind="port",
                sample="median of %d FP64 rhs! calls of the full workload (%d nodes) by the C oracle with "
                "OpenMP, %d threads, %s; rel err vs 1 thread %.1e" % (len(ts), N, nth, cpu, par_err),
                single_thread=single), du


def adjoint_ratio(L, configs):
    spec = configs.c4()
    ctx, info = setup_euler(L, spec, spec.side, visc=False)
    u0 = configs.perturbed_density(info["u0"], spec, info["N"])
    dt = info["dt"]
    ctx.upload(L.MF_BUF_U, u0)
    ctx.ssprk33(0.0, dt, 5, flags=L.MF_STEP_GRAPH)  # warm
    ctx.upload(L.MF_BUF_U, u0)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.ssprk33(0.0, dt, spec.nsteps)
    t_fwd = time.perf_counter() - t0
    ctx.upload(L.MF_BUF_U, u0)
    ctx.ssprk33_vjp(0.0, dt, spec.nsteps, spec.checkpoint_every)  # warm
    ctx.upload(L.MF_BUF_U, u0)
    ctx.synchronize()
    t0 = time.perf_counter()
    loss = ctx.ssprk33_vjp(0.0, dt, spec.nsteps, spec.checkpoint_every)
    t_adj = time.perf_counter() - t0
    return dict(nodes=info["N"], steps=spec.nsteps, checkpoint_every=spec.checkpoint_every,
                forward_ms=1e3 * t_fwd, adjoint_ms=1e3 * t_adj, ratio_total=t_adj / t_fwd,
                ratio_reverse=(t_adj - t_fwd) / t_fwd, loss=loss)


def _ensemble_runs(L, specs, steps):
    """One grouped context over the clouds (mf_set_points_grouped: every
    launch covers all of them, the steps replayed from its captured graph at
    the common time step), and for comparison the same clouds as separate
    contexts (one stream each, their own time steps) enqueued from this
    thread before any is synchronized."""
    from mfrbf.ensemble import Ensemble
    from mfrbf.equations import CompressibleEulerEquations2D

    t0 = time.perf_counter()
    ens = Ensemble(specs)
    t_setup_g = time.perf_counter() - t0
    g = ens.ctx
    g.upload(L.MF_BUF_U, ens.u0)
    g.ssprk33(0.0, ens.dt, 16, L.MF_PREC_SPLIT32, L.MF_STEP_GRAPH)  # capture, build the LDS tiles
    g.upload(L.MF_BUF_U, ens.u0)
    g.synchronize()
    w0 = time.perf_counter()
    g.ssprk33(0.0, ens.dt, steps, L.MF_PREC_SPLIT32, L.MF_STEP_GRAPH)
    wall_g = time.perf_counter() - w0
    ug = g.download(L.MF_BUF_U)
    eq = CompressibleEulerEquations2D(specs[0].gamma)
    err_g = 0.0
    for r, spec in enumerate(specs):
        lo, hi = ens.offsets[r], ens.offsets[r + 1]
        ex = spec.ic(np.c_[ens.x[lo:hi], ens.y[lo:hi]], steps * ens.dt, eq)
        err_g = max(err_g, float(np.abs(ug[0, lo:hi] - ex[0]).max()))
    rhs_ms, rhs_parts = g.time_rhs(steps * ens.dt, L.MF_PREC_SPLIT32, 50)
    total = ens.N
    evals = total * (3 * steps + 1)  # + the first (FSAL) rhs!
    out = dict(clouds=len(specs), nodes=total, steps=steps, dt=ens.dt, setup_s_grouped=t_setup_g,
               ms_grouped=1e3 * wall_g, node_evals_per_s=evals / wall_g,
               max_density_error_vs_exact=err_g,
               rhs_grouped={"ms": rhs_ms, "node_evals_per_s": total / (rhs_ms * 1e-3),
                            "kernels_ms": {"stencil": rhs_parts[0], "transpose": rhs_parts[1]}})
    # the split-FP32 stencil kernel with Flyer hyperviscosity, each array once:
    # u in, du out (FP64), 16-bit halo positions, (Dx, Dy) and H weights in FP32
    k = specs[0].k
    bpn = 4 * 8 + 4 * 8 + 2 * k + 8 * k + 4 * k
    out["rhs_grouped"]["bytes_per_node"] = bpn
    out["rhs_grouped"]["stencil_frac"] = bpn * total / (rhs_parts[0] * 1e-3) / 1e9 / HBM_PEAK_GBS
    del ens, g
    t0 = time.perf_counter()
    clouds = []
    for spec in specs:
        ctx, info = setup_euler(L, spec, spec.side, visc=False)
        ctx.compute_operator(spec.p, spec.deg, 2 * spec.hv["k"], L.MF_OP_HYPERVISCOSITY, want=False)
        ctx.set_source_hyperviscosity(spec.hv["c"] * info["dx_min"] ** (2 * spec.hv["k"]))
        clouds.append((spec, ctx, info))
    t_setup = time.perf_counter() - t0
    flags = L.MF_STEP_GRAPH | L.MF_STEP_NOSYNC

    def run():
        for _, ctx, info in clouds:
            ctx.upload(L.MF_BUF_U, info["u0"])
        w0 = time.perf_counter()
        for _, ctx, info in clouds:
            ctx.ssprk33(0.0, info["dt"], steps, L.MF_PREC_SPLIT32, flags)
        for _, ctx, _ in clouds:
            ctx.synchronize()
        return time.perf_counter() - w0

    run()  # capture the graphs, build the LDS tiles
    wall_c = run()
    out["contexts"] = dict(setup_s=t_setup, ms_concurrent=1e3 * wall_c, node_evals_per_s=evals / wall_c)
    del clouds
    # lanes: grouped contexts of at most 131,072 nodes, one stream each, at
    # the common time step (each cloud's state the grouped context's bits)
    from mfrbf.ensemble import EnsembleLanes
    t0 = time.perf_counter()
    lanes = EnsembleLanes(specs)
    t_setup_l = time.perf_counter() - t0

    def run_lanes():
        lanes.upload_initial()
        lanes.synchronize()
        w0 = time.perf_counter()
        lanes.ssprk33(0.0, steps)
        return time.perf_counter() - w0

    run_lanes()  # capture the graphs, build the LDS tiles
    wall_l = run_lanes()
    out["lanes"] = dict(lane_nodes=131072, lanes=len(lanes.lanes), setup_s=t_setup_l, ms=1e3 * wall_l,
                        node_evals_per_s=evals / wall_l)
    return out


def ensemble(L, configs, steps):
    """C5: 64 Euler clouds (16k..262k nodes) with Flyer hyperviscosity,
    split-FP32 rhs!, SSPRK33 steps, as one grouped context, as 64 contexts
    and as lanes (grouped contexts of at most 131,072 nodes on their own
    streams); and the same physics on 256 small clouds (48 x 48 nodes each),
    where the launches rather than the bytes set the time."""
    out = _ensemble_runs(L, configs.c5(), steps)
    small = configs.c5(256)
    for r, sp in enumerate(small):
        sp.name, sp.side = "small_%d" % r, 48
    out["small_clouds"] = _ensemble_runs(L, small, steps)
    return out


def advection_c1(L, configs, with_cpu):
    """C1, the reference's own CPU-sized case: linear advection of a Gaussian
    pulse on 4,096 nodes (r^3 + degree 2, 15-point stencils), 200 SSPRK33
    steps from one captured graph; with the C oracle's 200 steps on the same
    weights timed beside it and the final states compared."""
    from mfrbf.clouds import cloud_square
    from mfrbf.configs import advection_dt
    from mfrbf.equations import LinearScalarAdvectionEquation2D

    spec = configs.c1()
    x, y, tag, _, _ = cloud_square(spec.side, spec.seed)
    N = len(x)
    ctx = L.Context(0)
    ctx.set_equations(L.MF_EQ_ADVECTION_2D, spec.adv[0], spec.adv[1])
    ctx.set_points(x, y)
    idx, _, dx_min, _ = ctx.knn(spec.k)
    wx, wy = ctx.compute_operator(spec.p, spec.deg, 1, L.MF_OP_FLUX, want=True, k=spec.k)
    inflow = np.nonzero(tag == 1)[0].astype(np.int32)
    outflow = np.nonzero(tag == 2)[0].astype(np.int32)
    prm = spec.ic.params()
    ctx.set_boundary([0, len(inflow), len(inflow) + len(outflow)], [L.MF_BC_DIRICHLET, L.MF_BC_DO_NOTHING],
                     [spec.ic.code, spec.ic.code], np.concatenate([prm, prm]),
                     np.concatenate([inflow, outflow]), np.zeros((len(inflow) + len(outflow), 2)))
    u0 = np.ascontiguousarray(np.asarray(spec.ic(np.c_[x, y], 0.0, LinearScalarAdvectionEquation2D(spec.adv)),
                                         dtype=np.float64).reshape(1, N))
    dt, n = advection_dt(spec, dx_min), spec.nsteps
    ctx.upload(L.MF_BUF_U, u0)
    ctx.ssprk33(0.0, dt, n, flags=L.MF_STEP_GRAPH)  # warm: capture
    ts = []
    for _ in range(5):
        ctx.upload(L.MF_BUF_U, u0)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.ssprk33(0.0, dt, n, flags=L.MF_STEP_GRAPH)
        ctx.synchronize()
        ts.append(time.perf_counter() - t0)
    ug = ctx.download(L.MF_BUF_U)
    ms = 1e3 * float(np.median(ts))
    out = dict(nodes=N, steps=n, ms=ms, node_evals_per_s=3 * n * N / (ms * 1e-3),
               timing="median of 5 runs of 200 graph-replayed steps, host clock around a synchronized call")
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        prob = O.Problem(x, y, O.ell_ptr(N, spec.k), idx, wx, wy, O.EQ_ADVECTION, 1, adv=spec.adv,
                         boundary=[(O.BC_DIRICHLET, inflow, None, spec.ic.code, prm, None),
                                   (O.BC_NOTHING, outflow, None, None, None, None)])
        uo = u0.copy()
        t0 = time.perf_counter()
        prob.ssprk33(uo, 0.0, dt, n)
        el = time.perf_counter() - t0
        err = float(np.abs(ug - uo).max() / np.abs(uo).max())
        if err > 1e-12:
            raise RuntimeError("C1 final state differs from the oracle's: %.3e" % err)
        out["cpu_baseline"] = dict(value=3 * n * N / el, unit="node-evals/s", cores=1, kind="port",
                                   sample="the same 200 steps by the C oracle, 1 thread, %.2f s, %s"
                                   % (el, _cpu_model()))
        out["rel_err_vs_oracle"] = err
    return out


def weights_batch(L, configs, side, cpu_sample=0):
    from mfrbf.clouds import cloud_square

    sizes = configs.c3_stencils(side)
    x, y, _, _, _ = cloud_square(side, configs.SEED0 + 3)
    N = len(x)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    idx, _, _, _ = ctx.knn(int(sizes.max()))
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    mask = np.arange(idx.shape[1])[None, :] < sizes[:, None]
    ridx = np.ascontiguousarray(idx[mask]).astype(np.int32)
    ctx.weights_ragged_set(ptr, ridx)
    ctx.weights_ragged_run(5, 3, 1)
    ctx.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.weights_ragged_run(5, 3, 1)
    ctx.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / reps
    m = sizes + 10
    flops = float(np.sum(m ** 3 / 3.0 + 6.0 * m ** 2))
    out = dict(stencils=N, ms=ms, gflop=flops / 1e9, tflops=flops / (ms * 1e-3) / 1e12,
               frac_fp64_peak=flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS)
    if cpu_sample:
        out["cpu_baseline"] = weights_cpu(x, y, ptr, ridx, ctx.weights_ragged_get(), cpu_sample)
    return out


def weights_cpu(x, y, ptr, ridx, gpu_w, sample):
    """The C oracle's weights on the first `sample` stencils of the same
    batch (their ragged sizes as drawn): median of 3 timed passes on 1
    thread and on the host's share of cores (OpenMP), in stencils/s; the
    GPU's weights on that sample against the oracle's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    p = ptr[:sample + 1]
    nnz = int(p[-1])
    ridx = ridx[:nnz]
    nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    res = {}
    for name, threads in (("single_thread", None), ("openmp", nth)):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            wx, wy, rc = O.weights(x, y, p, ridx, 5, 3, 1, nthreads=threads)
            ts.append(time.perf_counter() - t0)
            if rc:
                raise RuntimeError("oracle weights: singular stencil %d" % (rc - 1))
        res[name] = dict(value=sample / float(np.median(ts)), unit="stencils/s", cores=threads or 1,
                         kind="port")
    err = max(float(np.abs(gpu_w[0][:nnz] - wx).max() / np.abs(wx).max()),
              float(np.abs(gpu_w[1][:nnz] - wy).max() / np.abs(wy).max()))
    res["sample"] = ("the first %d of the batch's stencils, FP64 C oracle, median of 3 passes each, %s"
                     % (sample, _cpu_model()))
    res["gpu_rel_err"] = err
    if err > 1e-9:
        raise RuntimeError("GPU weights differ from the oracle's: %.3e" % err)
    return res


# ------------------------------------------------------------ rocprofv3 summaries
# The profiles/ summaries are made by running rocprofv3 over this script
# and then this script's --profile-summary mode over the CSV files it wrote
# (the commands below, from the repository root on one MI355X, REPO = the
# repository root).
PROFILE_COMMANDS = {
    "traffic": [
        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 200 rocprofv3 --pmc FETCH_SIZE --output-format csv "
        "-d $REPO/gpurun_out/prof_fetch -o fetch -- python3 $REPO/bench.py --steps 10 --warmup 2 --no-extras --no-cpu",
        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 200 rocprofv3 --pmc WRITE_SIZE --output-format csv "
        "-d $REPO/gpurun_out/prof_write -o write -- python3 $REPO/bench.py --steps 10 --warmup 2 --no-extras --no-cpu",
        "python3 bench.py --profile-summary traffic gpurun_out/prof_fetch gpurun_out/prof_write",
    ],
    "trace": [
        "cd /tmp && export TMPDIR=/tmp && timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv "
        "-d $REPO/gpurun_out/prof_trace -o trace -- python3 $REPO/bench.py --steps 200 --warmup 10 --headline-only",
        "python3 bench.py --profile-summary trace gpurun_out/prof_trace",
    ],
    "weights": [
        "cd /tmp && export TMPDIR=/tmp && timeout -k 10 200 rocprofv3 --kernel-trace --stats --output-format csv "
        "-d $REPO/gpurun_out/wprof_trace -o trace -- python3 $REPO/bench.py --weights-only",
        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 120 rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_SALU "
        "SQ_INSTS_LDS SQ_WAVES SQ_WAVE_CYCLES SQ_INSTS_BRANCH SQ_INSTS_VALU_MFMA_F64 SQ_BUSY_CYCLES "
        "--output-format csv -d $REPO/gpurun_out/wprof_pmc1 -o p -- python3 $REPO/bench.py --weights-only",
        "python3 bench.py --profile-summary weights gpurun_out/wprof_trace gpurun_out/wprof_pmc1",
    ],
}
PROFILE_OUT = {"traffic": "rhs_euler_rv_262144_hbm_traffic", "trace": "rhs_euler_rv_262144_kernel_trace",
               "weights": "weights_c3_1048576_counters"}


def _csv_rows(d, suffix):
    import csv
    import glob

    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit("no *%s under %s" % (suffix, d))
    rows = []
    for fn in files:
        with open(fn, newline="") as f:
            rows += list(csv.DictReader(f))
    return rows


def _counter_medians(d):
    """{kernel: {counter: (median value per dispatch, dispatches)}} from the
    counter-collection CSVs under d."""
    per = {}
    for r in _csv_rows(d, "counter_collection.csv"):
        key = (r["Kernel_Name"], r.get("Dispatch_Id", r.get("Correlation_Id")))
        per.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], {})
        per[r["Kernel_Name"]][r["Counter_Name"]][key] = \
            per[r["Kernel_Name"]][r["Counter_Name"]].get(key, 0.0) + float(r["Counter_Value"])
    return {k: {c: (float(np.median(list(v.values()))), len(v)) for c, v in cs.items()} for k, cs in per.items()}


def _short(name, n=90):
    return name if len(name) <= n else name[:n]


def _trace_table(d, top=24):
    rows = _csv_rows(d, "kernel_stats.csv")
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | average ns | min ns | max ns | % |", "|---|---:|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append("| `%s` | %s | %.0f | %s | %s | %.2f |" % (_short(r["Name"]), r["Calls"], float(r["AverageNs"]),
                                                             r["MinNs"], r["MaxNs"], float(r["Percentage"])))
    return out, rows


def profile_summary(kind, dirs):
    """Writes profiles/<summary>.md (and .json for the traffic) from the
    rocprofv3 CSV output of the PROFILE_COMMANDS of `kind`."""
    stem = os.path.join(ROOT, "profiles", PROFILE_OUT[kind])
    cmds = "\n".join(PROFILE_COMMANDS[kind])
    if kind == "traffic":
        fetch, write = _counter_medians(dirs[0]), _counter_medians(dirs[1])
        # both calibration kernels move 16 bytes per double of a 262,144 x 4 state
        true_kib = 262144 * 4 * 8 / 1024.0
        cal = {}
        for name, tag in (("k_copy_buffer", "copy"), ("k_gather_probe", "gather")):
            kf = [k for k in fetch if name in k]
            kw = [k for k in write if name in k]
            if kf and kw:
                cal[tag] = (fetch[kf[0]]["FETCH_SIZE"][0], write[kw[0]]["WRITE_SIZE"][0])
        if "copy" not in cal:
            raise SystemExit("calibration kernel k_copy_buffer not in the profile")
        fs, ws = true_kib / cal["copy"][0], true_kib / cal["copy"][1]
        tab = {"fetch_scale": fs, "write_scale": ws, "calibration_kernel": "k_copy_buffer",
               "calibration_true_kib": true_kib, "kernels": {}}
        if "gather" in cal:
            tab["check_kernel"] = "k_gather_probe"
            tab["check_scaled_over_true"] = {"fetch": cal["gather"][0] * fs / true_kib,
                                             "write": cal["gather"][1] * ws / true_kib}
        lines = ["| kernel | dispatches | FETCH_SIZE KiB (median) | WRITE_SIZE KiB (median) | scaled MB per dispatch |",
                 "|---|---:|---:|---:|---:|"]
        for k in sorted(set(fetch) & set(write)):
            fk, nd = fetch[k]["FETCH_SIZE"]
            wk, _ = write[k]["WRITE_SIZE"]
            b = 1024.0 * (fk * fs + wk * ws)
            tab["kernels"][k] = {"fetch_kib": fk, "write_kib": wk, "bytes": b}
            lines.append("| `%s` | %d | %.1f | %.1f | %.1f |" % (_short(k), nd, fk, wk, b / 1e6))
        with open(stem + ".json", "w") as f:
            json.dump(tab, f, indent=1)
        chk = ""
        if "check_scaled_over_true" in tab:
            chk = ("The check kernel `k_gather_probe` (mf_bandwidth_probe kind 1: each wavefront gathers "
                   "scattered doubles within 2 KiB windows, the access shape of a stencil gather, %.0f KiB read "
                   "and written) then measures %.3f of its true reads and %.3f of its true writes after the "
                   "same scaling.\n\n" % (true_kib, tab["check_scaled_over_true"]["fetch"],
                                            tab["check_scaled_over_true"]["write"]))
        text = ("# HBM traffic per dispatch (FETCH_SIZE, WRITE_SIZE): Euler + RV rhs!, 262,144 nodes, k = 30\n\n"
                "Commands (from the repository root, on one MI355X):\n\n```\n%s\n```\n\n"
                "Calibration: `k_copy_buffer` (mf_bandwidth_probe kind 0, a plain copy) reads and writes exactly "
                "%.0f KiB; it reads FETCH_SIZE %.1f KiB and WRITE_SIZE %.1f KiB, so FETCH_SIZE is scaled by %.3f "
                "and WRITE_SIZE by %.3f.  %s%s\n" % (cmds, true_kib, cal["copy"][0], cal["copy"][1], fs, ws, chk,
                                                   "\n".join(lines)))
    elif kind == "trace":
        lines, _ = _trace_table(dirs[0])
        text = ("# Kernel trace: Euler + residual viscosity rhs!, 262,144 nodes, k = 30\n\n"
                "Commands (from the repository root, on one MI355X):\n\n```\n%s\n```\n\n"
                "Top kernels by total time.  The run includes the autotuner's sweep (every load-batch variant "
                "of each precision inside whole rhs! calls) and the headline (reference-order FP64) timing "
                "loop, so each kernel's average is over whole-rhs! sequences.\n\n%s\n" % (cmds, "\n".join(lines)))
    else:  # weights
        lines, rows = _trace_table(dirs[0], top=8)
        cnt = _counter_medians(dirs[1])
        names = ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_BRANCH", "SQ_INSTS_VALU_MFMA_F64",
                 "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES"]
        tl = ["| kernel | dispatches | waves | " + " | ".join(n[3:] + " per wave" for n in names[:6]) + " |",
              "|---|---:|---:|" + "---:|" * 6]
        for k, cs in sorted(cnt.items()):
            if "weights" not in k or "SQ_WAVES" not in cs:
                continue
            w = cs["SQ_WAVES"][0]
            tl.append("| `%s` | %d | %.0f | " % (_short(k), cs["SQ_WAVES"][1], w) +
                      " | ".join("%.0f" % (cs[n][0] / w) if n in cs else "-" for n in names[:6]) + " |")
        text = ("# Batched weights (C3): 1,048,576 ragged stencils (n = 30, 8%% with n in 40..50), PHS r^5 + "
                "degree 3\n\nCommands (from the repository root, on one MI355X):\n\n```\n%s\n```\n\n"
                "## Kernel trace\n\n%s\n\n## Instruction counters per wavefront (one wavefront per stencil; "
                "medians over dispatches)\n\n%s\n" % (cmds, "\n".join(lines), "\n".join(tl)))
    with open(stem + ".md", "w") as f:
        f.write(text)
    print(stem + ".md")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--side", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip adjoint and batched weights")
    ap.add_argument("--headline-only", action="store_true",
                    help="time the headline mode's rhs! only and print its figures (for profiled runs)")
    ap.add_argument("--weights-only", action="store_true",
                    help="time only the batched weights (1,048,576 ragged stencils) and print their JSON")
    ap.add_argument("--profile-summary", nargs="+", metavar=("KIND", "DIR"),
                    help="write profiles/ summaries from rocprofv3 CSV directories: traffic FETCH_DIR WRITE_DIR, "
                         "trace TRACE_DIR, or weights TRACE_DIR PMC_DIR (no GPU needed)")
    args = ap.parse_args()
    if args.profile_summary:
        profile_summary(args.profile_summary[0], args.profile_summary[1:])
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import mfrbf  # noqa: F401
    from mfrbf import _lib as L
    from mfrbf import configs, tune

    if args.weights_only:
        print(json.dumps(weights_batch(L, configs, 1024)), flush=True)
        return

    spec = configs.c2(args.side)
    log("setup: %d x %d cloud, k=%d" % (args.side, args.side, spec.k))
    ctx, info = setup_euler(L, spec, args.side)
    N, k = info["N"], spec.k
    ctx.upload(L.MF_BUF_U, info["u0"])
    # a plain copy of known size (N * 4 doubles read and written, three times):
    # the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters of a
    # profiled run of this script
    # and the windowed gather of the same size, the second calibration kernel
    # (both 16.8 MB per launch: launch latency weighs on them), and a
    # read-modify-write streamed over 512 MiB, the device's practical
    # streaming bandwidth
    probe = {"copy": [], "gather": [], "stream": []}
    for _ in range(3):
        ctx.copy(L.MF_BUF_AUX0, L.MF_BUF_U)
    for _ in range(5):
        probe["copy"].append(ctx.bandwidth_probe(0, L.MF_BUF_AUX0, L.MF_BUF_U))
        probe["gather"].append(ctx.bandwidth_probe(1, L.MF_BUF_AUX0, L.MF_BUF_U))
        probe["stream"].append(ctx.bandwidth_probe(2, L.MF_BUF_AUX0, L.MF_BUF_U))
    probe_bytes = {"copy": 16 * N * 4, "gather": 16 * N * 4, "stream": 2 * 8 * (1 << 26)}
    probe = {q: {"bytes": probe_bytes[q], "ms": float(np.min(v)),
                 "gbs": probe_bytes[q] / (np.min(v) * 1e-3) / 1e9} for q, v in probe.items()}
    # three SSPRK33 steps with the history callback so the residual viscosity is live
    ctx.ssprk33(0.0, info["dt"], 3, flags=L.MF_STEP_HISTORY)
    t_now = 3 * info["dt"]
    log("knn %.1f ms, weights %.1f ms; warmup %d rhs!" % (1e3 * info["t_knn"], 1e3 * info["t_weights"],
                                                          args.warmup))
    ctx.time_rhs(t_now, L.MF_PREC_F64_REF, max(1, args.warmup))
    ctx.synchronize()
    log("autotune: load-batch sizes of the stencil and transposed kernels")
    tuning = tune.autotune(ctx, t_now, N, k, L.MF_EQ_EULER_2D, L.MF_VISC_RESIDUAL,
                           precisions=(L.MF_PREC_F64, L.MF_PREC_SPLIT32, L.MF_PREC_F64_REF),
                           use_cache=False)
    log("tuning: " + json.dumps({p: (v["flux"], v["transpose"]) for p, v in tuning.items()}))
    # the headline FP64 mode is the reference's operation order (the one within
    # 1e-12 of the reference at this size); the fused-multiply-add FP64 and
    # the split-FP32 modes are timed beside it
    log("timing %d rhs! (FP64, reference operation order)" % args.steps)
    w0 = time.perf_counter()
    rhs_ms, parts = ctx.time_rhs(t_now, L.MF_PREC_F64_REF, args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - w0
    value = N / (rhs_ms * 1e-3)
    tinfo = ctx.tile_info()
    ba, bb = kernel_bytes(N, k, tinfo)
    kernels = {"stencil": (parts[0], ba), "transpose": (parts[1], bb)}
    dom = max(kernels, key=lambda q: kernels[q][0])
    kms, kbytes = kernels[dom]
    achieved = kbytes / (kms * 1e-3) / 1e9
    rbytes = rhs_bytes(N, k)
    traffic, traffic_kernel = measured_traffic(dom, tuning, args.side,
                                               tinfo["tiles_b"] if dom == "transpose" else tinfo["tiles_a"])
    log("rhs %.4f ms (%.3e node-evals/s); stencil %.4f ms, transpose %.4f ms, reductions %.4f ms, bc %.4f ms"
        % (rhs_ms, value, parts[0], parts[1], parts[2], parts[3]))
    if args.headline_only:
        # for a profiled run: the headline mode's calls only
        print(json.dumps({"metric": METRIC, "value": value, "unit": "node-evals/s", "ms_per_step": rhs_ms,
                          "kernels_ms": {"stencil": parts[0], "transpose": parts[1], "reductions": parts[2]}}),
              flush=True)
        return
    log("timing %d rhs! (FP64 fused multiply-add, split-FP32)" % args.steps)
    f_ms, f_parts = ctx.time_rhs(t_now, L.MF_PREC_F64, args.steps)
    s_ms, s_parts = ctx.time_rhs(t_now, L.MF_PREC_SPLIT32, args.steps)
    fa, fb = kernel_bytes(N, k, tinfo)
    log("fma rhs %.4f ms; split-FP32 rhs %.4f ms" % (f_ms, s_ms))
    # the same call with the neighbour gathers from global memory (no LDS tiles)
    ctx.set_tuning(L.MF_TUNE_TILES, 0)
    nt_ms, nt_parts = ctx.time_rhs(t_now, L.MF_PREC_F64_REF, max(200, args.steps // 10))
    ctx.set_tuning(L.MF_TUNE_TILES, 1)
    ctx.time_rhs(t_now, L.MF_PREC_F64_REF, 3)
    log("without LDS tiles: rhs %.4f ms; stencil %.4f ms, transpose %.4f ms" % (nt_ms, nt_parts[0], nt_parts[1]))
    # each call after a 512 MiB read-modify-write that evicts the L2 and
    # last-level caches (median of 20 eager calls)
    cold = {name: ctx.time_rhs_cold(t_now, prec, 20)
            for name, prec in (("f64_ref", L.MF_PREC_F64_REF), ("split32", L.MF_PREC_SPLIT32))}
    log("cold caches: f64_ref %.4f ms, split32 %.4f ms" % (cold["f64_ref"], cold["split32"]))
    dus = {}
    for name, prec in (("f64_ref", L.MF_PREC_F64_REF), ("f64_fma", L.MF_PREC_F64), ("split32", L.MF_PREC_SPLIT32)):
        ctx.rhs(t_now, prec)
        dus[name] = ctx.download(L.MF_BUF_DU)
    du64 = dus["f64_ref"]
    s_err = float(np.abs(dus["split32"] - du64).max() / np.abs(du64).max())
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "node-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rhs_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Euler + residual viscosity rhs!, %d jittered nodes (%dx%d), 30-point "
                               "PHS r^5 + degree-3 stencils, FP64 in the reference's operation order"
                               % (N, args.side, args.side),
                   "nodes": N, "stencil": k, "precision": "f64_ref", "timed_wall_s": wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "traffic_kernel": traffic_kernel,
                     "kernel_ms": kms, "kernel_bytes": kbytes,
                     "rhs_frac": rbytes / (rhs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "rhs_bytes": rbytes,
                     "rhs_frac_cold": rbytes / (cold["f64_ref"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "probe_gbs": {q: v["gbs"] for q, v in probe.items()},
                     "frac_of_stream": achieved / probe["stream"]["gbs"]},
        "kernels_ms": {"stencil": parts[0], "transpose": parts[1], "reductions": parts[2],
                       "boundary": parts[3]},
        "f64_fma": {"value": N / (f_ms * 1e-3), "ms_per_step": f_ms,
                    "kernels_ms": {"stencil": f_parts[0], "transpose": f_parts[1], "reductions": f_parts[2]},
                    "stencil_frac": fa / (f_parts[0] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "transpose_frac": fb / (f_parts[1] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "rhs_frac": rbytes / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "split32": {"value": N / (s_ms * 1e-3), "ms_per_step": s_ms, "speedup_vs_f64": rhs_ms / s_ms,
                    "speedup_vs_f64_fma": f_ms / s_ms, "rel_err_vs_f64": s_err,
                    "kernels_ms": {"stencil": s_parts[0], "transpose": s_parts[1]}},
        "cold_cache": {"ms_per_step": cold, "node_evals_per_s": {q: N / (v * 1e-3) for q, v in cold.items()},
                       "flush": "512 MiB read-modify-write before each of 20 eager calls, median"},
        "lds_tiles": tinfo,
        "f64_global_gathers": {"ms_per_step": nt_ms,
                               "kernels_ms": {"stencil": nt_parts[0], "transpose": nt_parts[1]}},
        "setup_ms": {"knn": 1e3 * info["t_knn"], "weights": 1e3 * info["t_weights"]},
        "tuning": tuning,
    }
    if not args.no_extras:
        log("adjoint: 100 SSPRK33 steps, 65,536 nodes")
        out["adjoint"] = adjoint_ratio(L, configs)
        log("C1: advection, 4,096 nodes, 200 SSPRK33 steps")
        out["c1_advection"] = advection_c1(L, configs, not args.no_cpu)
        log("batched weights: 1,048,576 ragged stencils (and the CPU oracle on a sample)")
        out["weights"] = weights_batch(L, configs, 1024, 0 if args.no_cpu else 32768)
        log("ensemble: 64 clouds, 50 SSPRK33 steps each (split-FP32; one grouped context, and 64 contexts)")
        out["ensemble"] = ensemble(L, configs, 50)
        log("Tominec hyperviscosity after the residual viscosity, %d nodes" % N)
        out["tominec"] = tominec(L, ctx, info, spec, t_now, max(200, args.steps // 4))
        log("IGR source after the residual viscosity, %d nodes" % N)
        out["igr"] = igr(L, ctx, info, spec, t_now, 50)
        ctx.time_rhs(t_now, L.MF_PREC_F64_REF, 3)
    if not args.no_cpu:
        log("cpu baseline (oracle, %.0f s budget)" % args.cpu_seconds)
        cb, du_cpu = cpu_baseline(L, ctx, info, spec, t_now, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["rel_err_vs_oracle"] = {name: float(np.abs(d - du_cpu).max() / np.abs(du_cpu).max())
                                    for name, d in dus.items()}
    if not args.no_extras:
        # last: it advances the state and the history the comparisons above use
        log("SSPRK33 steps with the history callback and the stage limiter, %d nodes" % N)
        out["stepping"] = stepping(L, ctx, info, 64)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
