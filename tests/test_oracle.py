"""Pins of the CPU oracle (no GPU): the reference's own known-answer and
identity tests, library routines, and what the mathematics fixes."""
import os
import sys

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
from scipy.spatial import cKDTree

import cases
import oracle as O
from conftest import GOLDEN, rel_inf
from mfrbf import configs
from mfrbf.clouds import cloud_square
from mfrbf.domain import read_medusa_file


# ------------------------------------------------------------------ clouds
@pytest.mark.parametrize("s,seed", [(64, 20261202), (101, 3), (512, 20261203)])
def test_cloud_generator_c_equals_numpy(s, seed):
    a = O.cloud_square(s, seed)
    b = cloud_square(s, seed)
    for p, q in zip(a, b):
        assert np.array_equal(p, q)


def test_cloud_properties():
    s = 64
    x, y, tag, nx, ny = O.cloud_square(s, 11)
    h = 1.0 / (s - 1)
    i, j = np.tile(np.arange(s), s), np.repeat(np.arange(s), s)
    assert np.all(np.abs(x - i * h) <= 0.3 * h + 1e-15) and np.all(np.abs(y - j * h) <= 0.3 * h + 1e-15)
    corners = [0, s - 1, s * (s - 1), s * s - 1]
    assert np.all(np.isin(x[corners], [0.0, 1.0])) and np.all(np.isin(y[corners], [0.0, 1.0]))
    assert np.all(x[i == 0] == 0.0) and np.all(y[j == 0] == 0.0)
    assert np.all(tag[(i == 0) | (j == 0)] == 1) and np.sum(tag == 2) == 2 * (s - 1) - 1
    assert np.allclose(np.hypot(nx[tag > 0], ny[tag > 0]), 1.0)


# --------------------------------------------------------------------- kNN
def test_knn_brute_matches_ckdtree_and_grid():
    x, y, _, _, _ = O.cloud_square(48, 20261202)
    ib, d2 = O.knn_brute(x, y, 30)
    dd, ii = cKDTree(np.c_[x, y]).query(np.c_[x, y], k=30)
    assert np.array_equal(ib, ii)
    assert np.allclose(np.sqrt(d2), dd, rtol=1e-15, atol=1e-16)
    ig, dg = O.knn_grid(x, y, 30)
    assert np.array_equal(ib, ig) and np.array_equal(d2, dg)
    assert np.all(ib[:, 0] == np.arange(len(x)))
    assert np.all(np.diff(d2, axis=1) >= 0)


def test_knn_tie_order_by_index():
    # exact lattice: ties everywhere, resolved by index
    s = 9
    x = np.tile(np.arange(s), s).astype(float)
    y = np.repeat(np.arange(s), s).astype(float)
    ib, d2 = O.knn_brute(x, y, 9)
    for i in range(len(x)):
        for a in range(8):
            assert (d2[i, a], ib[i, a]) < (d2[i, a + 1], ib[i, a + 1])
    ig, _ = O.knn_grid(x, y, 9)
    assert np.array_equal(ib, ig)


def test_knn_grid_degenerate_extents():
    """The grid search on flat (collinear along either axis), nearly flat
    and elongated clouds equals brute force: its grid stays at most N + 1
    cells along the long side instead of a cell count that overflows."""
    rng = np.random.default_rng(20261210)
    n = 600
    t = rng.random(n)
    for x, y in ((t, np.full(n, 0.25)), (np.full(n, 0.7), t), (t, 0.5 + 1e-13 * rng.standard_normal(n)),
                 (1000.0 * t, rng.random(n))):
        ib, d2 = O.knn_brute(x, y, 15)
        ig, dg = O.knn_grid(x, y, 15)
        assert np.array_equal(ib, ig) and np.array_equal(d2, dg)


def test_knn_grid_large_vs_ckdtree():
    x, y, _, _, _ = O.cloud_square(256, 20261205)
    ig, dg = O.knn_grid(x, y, 30)
    dd, ii = cKDTree(np.c_[x, y]).query(np.c_[x, y], k=31)
    # no near-ties across the stencil, nor at the k / k+1 boundary
    gaps = np.diff(dd, axis=1)
    assert gaps[:, 1:].min() > 1e-12
    assert np.array_equal(ig, ii[:, :30])


def _config_clouds():
    out = [("C1", 64, configs.SEED0 + 1, 15), ("C2", 512, configs.SEED0 + 2, 30),
           ("C3", 1024, configs.SEED0 + 3, 50), ("C4", 256, configs.SEED0 + 4, 30)]
    for r, spec in enumerate(configs.c5()):
        out.append(("C5-%d" % r, spec.side, spec.seed, spec.k))
    return out


def test_config_clouds_have_no_near_ties():
    """Every configuration's cloud (C1-C4 and the 64 C5 clouds, C3 with its
    largest stencil k = 50): along each node's k + 1 nearest neighbours
    consecutive distances differ by more than 1e-12 of the distance (about
    4,500 units in the last place; the closest pair, in C5 cloud 25, is
    1.3e-10 apart), so the stencils, their order and their membership at k
    do not depend on how a kNN search breaks ties or rounds distances; the k
    nearest by scipy's cKDTree equal the oracle's grid search."""
    for name, side, seed, k in _config_clouds():
        x, y, _, _, _ = O.cloud_square(side, seed)
        dd, ii = cKDTree(np.c_[x, y]).query(np.c_[x, y], k=k + 1, workers=-1)
        rel_gap = (np.diff(dd, axis=1)[:, 1:] / dd[:, 2:]).min()
        assert rel_gap > 1e-12, (name, float(rel_gap))
        if side <= 512:
            ig, _ = O.knn_grid(x, y, k)
            assert np.array_equal(ig, ii[:, :k]), name


def test_dx_stats():
    x, y, _, _, _ = O.cloud_square(64, 20261202)
    _, d2 = O.knn_brute(x, y, 2)
    mn, av = O.dx_stats(d2[:, 1])
    dd, _ = cKDTree(np.c_[x, y]).query(np.c_[x, y], k=2)
    assert mn == dd[:, 1].min()
    assert abs(av - dd[:, 1].mean()) < 4e-15 * av
    # Julia's pairwise reduction: exact on integers
    assert O.pairwise_sum(np.arange(5000.0)) == 5000 * 4999 / 2


# ------------------------------------------------------------ Bunch-Kaufman
def _saddle(m, np_, rng):
    n = m - np_
    A = rng.standard_normal((m, m))
    A = A + A.T
    A[n:, n:] = 0.0
    return A


@pytest.mark.parametrize("m", [6, 21, 40, 60])
def test_bk_solve_matches_lapack_sysv(m):
    rng = np.random.default_rng(m)
    A = _saddle(m, 3 if m < 10 else (6 if m < 40 else 10), rng)
    B = rng.standard_normal((m, 2))
    X, info = O.bk_solve_full(A, B)
    assert info == 0
    Xs = sla.solve(A, B, assume_a="sym")
    assert rel_inf(X, Xs) < 1e-11
    assert rel_inf(A @ X, B) < 1e-11


def test_bk_factor_matches_lapack_sytrf():
    rng = np.random.default_rng(7)
    for m in (12, 17, 33):
        A = _saddle(m, 5, rng)
        F, ipiv, info = O.bk_factor(A)
        lu, piv, _ = sla.lapack.dsytrf(A, lower=0)
        assert np.array_equal(ipiv, piv)  # same 1x1 / 2x2 pivot sequence
        assert rel_inf(np.triu(F), np.triu(lu)) < 1e-12


def test_bk_two_by_two_pivots_and_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    A = np.array([[0.0, 1.0, 2.0, 0.5], [1.0, 0.0, 3.0, 1.0], [2.0, 3.0, 0.0, 1.0], [0.5, 1.0, 1.0, 0.0]])
    b = np.array([[1.0], [2.0], [3.0], [4.0]])
    F, ipiv, info = O.bk_factor(A)
    assert np.any(ipiv < 0)  # zero diagonal forces 2x2 pivots
    X, _ = O.bk_solve_full(A, b)
    Xm = mpmath.lu_solve(mpmath.matrix(A.tolist()), mpmath.matrix(b.tolist()))
    assert np.allclose(X.ravel(), [float(v) for v in Xm], rtol=1e-14)


# ----------------------------------------------------------------- weights
def test_phs_derivatives_vs_sympy():
    sympy = pytest.importorskip("sympy")
    xs, ys = sympy.symbols("x y", real=True)
    rng = np.random.default_rng(3)
    for p in (3, 5):
        phi = (xs ** 2 + ys ** 2) ** sympy.Rational(p, 2)
        for k in (0, 1, 2, 4):
            fx = sympy.lambdify((xs, ys), sympy.diff(phi, xs, k), "mpmath")
            fy = sympy.lambdify((xs, ys), sympy.diff(phi, ys, k), "mpmath")
            for _ in range(5):
                a, b = rng.uniform(-1, 1, 2)
                assert O.phs_deriv(p, k, 0, a, b) == pytest.approx(float(fx(a, b)), rel=1e-13, abs=1e-14)
                assert O.phs_deriv(p, k, 1, a, b) == pytest.approx(float(fy(a, b)), rel=1e-13, abs=1e-14)


@pytest.mark.parametrize("p,deg,k", [(3, 2, 15), (5, 3, 30), (3, 3, 20)])
def test_weights_reproduce_polynomial_derivatives(p, deg, k):
    x, y, _, _, _ = O.cloud_square(40, 20261202)
    N = len(x)
    idx, _ = O.knn_grid(x, y, k)
    wx, wy, rc = O.weights(x, y, O.ell_ptr(N, k), idx, p, deg, 1)
    assert rc == 0
    wx, wy = wx.reshape(N, k), wy.reshape(N, k)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            f = x ** a * y ** b
            fx = a * x ** max(a - 1, 0) * y ** b if a else 0 * x
            fy = b * x ** a * y ** max(b - 1, 0) if b else 0 * x
            assert np.abs((wx * f[idx]).sum(1) - fx).max() < 1e-9
            assert np.abs((wy * f[idx]).sum(1) - fy).max() < 1e-9
    assert np.abs(wx.sum(1)).max() < 1e-10 * np.abs(wx).max()
    # a missing scale factor, sign or transposed rhs column fails the line above;
    # an independent dense solve of the same saddle system agrees
    e = 123
    X = np.c_[x[idx[e]] - x[idx[e, 0]], y[idx[e]] - y[idx[e, 0]]]
    s = 1.0 / np.abs(X).max(0)
    Xs = X * s
    R = np.linalg.norm(Xs[:, None] - Xs[None, :], axis=2) ** p
    mons = [(d - t, t) for d in range(deg + 1) for t in range(d + 1)]
    P = np.stack([Xs[:, 0] ** ax * Xs[:, 1] ** ay for ax, ay in mons], 1)
    M = np.block([[R, P], [P.T, np.zeros((len(mons), len(mons)))]])
    mir = -Xs.copy()
    mir[0] = np.finfo(float).eps
    r = np.linalg.norm(mir, axis=1)
    rhs = np.r_[p * mir[:, 0] * r ** (p - 2), [1.0 if m == (1, 0) else 0.0 for m in mons]]
    w = sla.solve(M, rhs, assume_a="sym")[:k] * s[0]
    assert rel_inf(wx[e], w) < 1e-10


def test_weights_par_matches_serial():
    """The CPU baseline's OpenMP weights (chunks of 64 ragged stencils) give
    the serial loop's bits."""
    x, y, _, _, _ = O.cloud_square(30, 20261203)
    N = len(x)
    idx, _ = O.knn_grid(x, y, 45)
    sizes = np.where(np.arange(N) % 7 == 0, 45, 30)  # ragged, two chunk-crossing sizes
    ptr = np.r_[0, np.cumsum(sizes)].astype(np.int64)
    ridx = np.concatenate([idx[e, :sizes[e]] for e in range(N)])
    ws = O.weights(x, y, ptr, ridx, 5, 3, 1)
    wp = O.weights(x, y, ptr, ridx, 5, 3, 1, nthreads=3)
    assert ws[2] == 0 and wp[2] == 0
    assert np.array_equal(ws[0], wp[0]) and np.array_equal(ws[1], wp[1])


def test_hyperviscosity_weights_annihilate_cubics():
    x, y, _, _, _ = O.cloud_square(40, 20261202)
    N, k = len(x), 30
    idx, _ = O.knn_grid(x, y, k)
    hx, hy, rc = O.weights(x, y, O.ell_ptr(N, k), idx, 5, 3, 4)
    assert rc == 0
    H = (hx + hy).reshape(N, k)
    scale = np.abs(H).max()
    for a in range(4):
        for b in range(4 - a):
            f = x ** a * y ** b
            assert np.abs((H * f[idx]).sum(1)).max() < 1e-8 * scale


# -------------------------------------------------------------- operators
def test_csc_products_match_scipy():
    x, y, _, _, _ = O.cloud_square(32, 1)
    N, k = len(x), 15
    idx, _ = O.knn_grid(x, y, k)
    w = np.random.default_rng(0).standard_normal((N, k))
    A = sp.csr_matrix((w.ravel(), idx.ravel(), np.arange(N + 1) * k), shape=(N, N)).tocsc()
    csc = O.CSC(N, O.ell_ptr(N, k), idx)
    nz = csc.values(w)
    v = np.random.default_rng(1).standard_normal(N)
    assert rel_inf(csc.mul(nz, v), A @ v) < 1e-14
    assert rel_inf(csc.tmul(nz, v), A.T @ v) < 1e-14
    assert np.array_equal(csc.colptr, A.indptr) and np.array_equal(csc.rowval, A.indices)


def test_euler_flux_textbook():
    g = 1.4
    rng = np.random.default_rng(2)
    for _ in range(10):
        rho, v1, v2, p = rng.uniform(0.5, 2, 4)
        E = p / (g - 1) + 0.5 * rho * (v1 * v1 + v2 * v2)
        u = np.array([rho, rho * v1, rho * v2, E])
        f1 = O.flux(O.EQ_EULER, u, 0, g)
        f2 = O.flux(O.EQ_EULER, u, 1, g)
        assert np.allclose(f1, [rho * v1, rho * v1 * v1 + p, rho * v1 * v2, (E + p) * v1], rtol=1e-14)
        assert np.allclose(f2, [rho * v2, rho * v1 * v2, rho * v2 * v2 + p, (E + p) * v2], rtol=1e-14)
    assert O.flux(O.EQ_ADVECTION, np.array([2.0]), 1, adv=(0.5, -3.0))[0] == -6.0


# ------------------------------------------------------------------- rhs!
def _cyl():
    return read_medusa_file(os.path.join(GOLDEN, "cyl_0_05", "cyl_0_05"))


def test_medusa_fixture_groups():
    pos, interior, bidx, bnorm, normals, ranges, types = _cyl()
    assert pos.shape == (2154, 2) and len(interior) == 1928
    assert [len(g) for g in bidx] == [40, 40, 60, 60, 26]


def _cyl_problem(p=3, deg=3, bcs=None, rv_mode=O.RV_NONE):
    pos, _, bidx, bnorm, _, _, _ = _cyl()
    x, y = pos[:, 0].copy(), pos[:, 1].copy()
    N = len(x)
    k = 20  # PointCloudBasis(Point2D(), 3) stencil size
    idx, _ = O.knn_brute(x, y, k)
    wx, wy, rc = O.weights(x, y, O.ell_ptr(N, k), idx, p, deg, 1)
    assert rc == 0
    _, d2 = O.knn_brute(x, y, 2)
    dmin, davg = O.dx_stats(d2[:, 1])
    bcs = bcs or []
    boundary = [(kind, bidx[g - 1], bnorm[g - 1], fn, None, None) for (kind, g, fn) in bcs]
    prob = O.Problem(x, y, O.ell_ptr(N, k), idx, wx, wy, O.EQ_EULER, 4, boundary=boundary, rv_mode=rv_mode,
                     dx_avg=davg)
    u0 = np.stack([1.4 + 0.1 * x + 0.1 * y, 4.1 + 0.1 * x + 0.1 * y, 0.0 + 0.1 * x + 0.1 * y,
                   8.8 + 0.1 * x + 0.1 * y])
    return prob, np.ascontiguousarray(u0), idx, wx.reshape(N, k), wy.reshape(N, k)


def test_divergence_identity_cyl():
    """test/divergence_test.jl:63-82: calc_fluxes! == -Dx*Fx - Dy*Fy."""
    prob, u0, idx, wx, wy = _cyl_problem()
    N = u0.shape[1]
    du, _, _ = prob.rhs(u0.copy(), 0.0)
    Dx = sp.csr_matrix((wx.ravel(), idx.ravel(), np.arange(N + 1) * 20), shape=(N, N))
    Dy = sp.csr_matrix((wy.ravel(), idx.ravel(), np.arange(N + 1) * 20), shape=(N, N))
    F1 = np.stack([O.flux(O.EQ_EULER, u0[:, e], 0) for e in range(N)], 1)
    F2 = np.stack([O.flux(O.EQ_EULER, u0[:, e], 1) for e in range(N)], 1)
    ref = -(Dx @ F1.T).T - (Dy @ F2.T).T
    assert rel_inf(du, ref) < 1e-12


def test_upwind_viscosity_identity_cyl():
    """test/upwind_viscosity_test.jl:67-78: UV source == -Dx'(eps Dx u) - Dy'(eps Dy u)."""
    prob, u0, idx, wx, wy = _cyl_problem(rv_mode=O.RV_UPWIND)
    N = u0.shape[1]
    du, eps, code = prob.rhs(u0.copy(), 0.0)
    prob0, *_ = _cyl_problem()
    du0, _, _ = prob0.rhs(u0.copy(), 0.0)
    Dx = sp.csr_matrix((wx.ravel(), idx.ravel(), np.arange(N + 1) * 20), shape=(N, N))
    Dy = sp.csr_matrix((wy.ravel(), idx.ravel(), np.arange(N + 1) * 20), shape=(N, N))
    rho, m1, m2, E = u0
    v1, v2 = m1 / rho, m2 / rho
    p = 0.4 * (E - 0.5 * (m1 * v1 + m2 * v2))
    eps_ref = 0.5 * prob.cfg.dx_avg * (np.hypot(v1, v2) + np.sqrt(1.4 * p / rho))
    assert rel_inf(eps, eps_ref) < 1e-14 and np.all(code == 1)
    src = -(Dx.T @ (eps_ref[:, None] * (Dx @ u0.T))).T - (Dy.T @ (eps_ref[:, None] * (Dy @ u0.T))).T
    assert rel_inf(du - du0, src) < 1e-9


def test_boundary_conditions_cyl():
    prob, u0, idx, wx, wy = _cyl_problem(bcs=[(O.BC_DIRICHLET, 1, O.FN_GRADIENT), (O.BC_NOTHING, 2, None),
                                              (O.BC_SLIP, 3, None), (O.BC_SLIP, 4, None), (O.BC_SLIP, 5, None)])
    pos, _, bidx, bnorm, _, _, _ = _cyl()
    u = u0.copy()
    u[:, bidx[0]] += 1.0  # Dirichlet must restore the gradient state
    du, _, _ = prob.rhs(u, 0.0)
    assert np.allclose(u[:, bidx[0]], u0[:, bidx[0]], rtol=0, atol=1e-15)
    assert np.all(du[:, bidx[0]] == 0.0)
    for g in (2, 3, 4):
        n = bnorm[g] / np.linalg.norm(bnorm[g], axis=1)[:, None]
        mom = u[1:3, bidx[g]].T
        assert np.abs((mom * n).sum(1)).max() < 1e-14
        assert np.all(du[1:3, bidx[g]] == 0.0)
    # do-nothing: the outlet keeps its state and its flux divergence
    assert np.array_equal(u[:, bidx[1]], u0[:, bidx[1]])
    assert np.all(np.any(du[:, bidx[1]] != 0.0, axis=0))


def test_free_stream_and_linear_exactness():
    spec = configs.c1()
    c = cases.build(spec, side=32, use_gpu=False)
    # linear field advected: du = -(a1 + a2) exactly up to roundoff (no boundary conditions)
    x, y = c.x, c.y
    prob = O.Problem(x, y, c.ptr, c.idx, c.wx, c.wy, O.EQ_ADVECTION, 1, adv=(1.0, 1.0))
    du, _, _ = prob.rhs(np.ascontiguousarray((2.0 * x + 3.0 * y)[None, :]), 0.0)
    assert np.abs(du + 5.0).max() < 1e-10
    du, _, _ = prob.rhs(np.ones((1, len(x))), 0.0)
    assert np.abs(du).max() < 1e-10


# ------------------------------------------------------ residual viscosity
def test_rv_norms_quirks():
    # mean divides by 4N; max is lexicographic over |u - mean|
    u = np.array([[1.0, 3.0], [10.0, 0.0], [0.0, 0.0], [2.0, 2.0]])
    mean, ninf = O.rv_norms(u)
    assert np.array_equal(mean, np.array([4.0, 10.0, 0.0, 4.0]) / 8.0)
    d = np.abs(u - mean[:, None])
    j = 1 if tuple(d[:, 1]) > tuple(d[:, 0]) else 0
    expect = d[:, j].copy()
    expect[expect == 0.0] = np.finfo(float).eps
    assert np.array_equal(ninf, expect)
    assert not np.array_equal(ninf, np.maximum(d[:, 0], d[:, 1]))  # differs from element-wise


def test_update_visc_branches():
    import ctypes as C
    L = O.lib()
    uw = np.array([1.0, 1.0, np.nan, 2.0, 3.0])
    rv = np.array([0.5, np.inf, 0.1, 2.0, 1.0])
    eps, code = np.zeros(5), np.zeros(5, np.int8)
    L.or_update_visc(5, O.d(uw), O.d(rv), 1, O.d(eps), code.ctypes.data_as(C.POINTER(C.c_int8)))
    # min(0.1, NaN) is NaN (Julia's min propagates NaN), choice code 1
    assert list(code) == [0, 1, 1, 1, 0]
    assert eps[0] == 0.5 and eps[1] == 1.0 and np.isnan(eps[2]) and eps[3] == 2.0 and eps[4] == 1.0
    L.or_update_visc(5, O.d(uw), O.d(rv), 0, O.d(eps), code.ctypes.data_as(C.POINTER(C.c_int8)))
    assert list(code) == [1, 1, 2, 1, 1] and eps[2] == np.finfo(float).eps


# Known answers for the residual-viscosity indicator of a whole rhs!
# (hyperviscosity.jl:289-349, auxiliary/mpi.jl:40-52, 71-82), computed here
# from the formulas with numpy and compared with the oracle's eps.  Each
# test also evaluates the plausible mistakes (h for h^2, dx_min for dx_avg,
# mean over N instead of 4N, component-wise max, du after the second
# boundary pass, u before the first) and checks that every one of them
# would miss the oracle by far more than the tolerance.
def _euler_flux_np(u, gamma=1.4):
    rho, m1, m2, E = u
    v1, v2 = m1 / rho, m2 / rho
    p = (gamma - 1.0) * (E - 0.5 * (m1 * v1 + m2 * v2))
    f1 = np.stack([m1, m1 * v1 + p, m2 * v1, (E + p) * v1])
    f2 = np.stack([m2, m1 * v2, m2 * v2 + p, (E + p) * v2])
    return f1, f2, v1, v2, p


def _eps_rv_np(c, u_seen, du_seen, approx_du, h, c_rv, mean_div=4, lex=True):
    N = u_seen.shape[1]
    mean = u_seen.sum(axis=1) / (mean_div * N)
    dev = np.abs(u_seen - mean[:, None])
    if lex:
        j = max(range(N), key=lambda i: tuple(dev[:, i]))
        ninf = dev[:, j].copy()
    else:
        ninf = dev.max(axis=1)
    ninf[ninf == 0.0] = np.finfo(float).eps
    return 0.5 * c_rv * h * h * (np.abs(approx_du - du_seen) / ninf[:, None]).max(axis=0)


def test_residual_viscosity_constant_state_closed_form():
    """u = c at every node (no boundary conditions): the flux divergence
    vanishes to rounding, the mean is c/4 (sum over all 4N scalars), every
    node ties at |c - c/4| = 3|c|/4, so n_v = 3|c_v|/4 and
    eps_rv,i = c_rv/2 * dx_avg^2 * max_v |approx_du_v,i| / (3|c_v|/4)."""
    c = cases.build(configs.c2(), side=14, use_gpu=False, bc=("nothing", "nothing"))
    N = c.N
    const = np.array([1.0, 0.5, -0.25, 2.5])
    u = np.ascontiguousarray(np.repeat(const[:, None], N, axis=1))
    rng = np.random.default_rng(3)
    A = rng.uniform(0.5, 1.5, (4, N)) * rng.choice([-1.0, 1.0], (4, N))
    c_rv, h = 0.7, c.dx_avg
    assert abs(c.dx_min / c.dx_avg - 1.0) > 0.05  # the two spacings are distinguishable
    prob = O.Problem(c.x, c.y, c.ptr, c.idx, c.wx, c.wy, O.EQ_EULER, 4, rv_mode=O.RV_RESIDUAL, c_rv=c_rv,
                     c_uw=1e6, dx_avg=h)
    _, eps, code = prob.rhs(u.copy(), 0.0, success_iter=2, approx_du=A)
    ninf = 0.75 * np.abs(const)
    expect = 0.5 * c_rv * h * h * (np.abs(A) / ninf[:, None]).max(axis=0)
    assert np.all(code == 0)
    assert rel_inf(eps, expect) < 1e-10
    wrong = {
        "h": 0.5 * c_rv * h * (np.abs(A) / ninf[:, None]).max(axis=0),
        "dx_min": 0.5 * c_rv * c.dx_min ** 2 * (np.abs(A) / ninf[:, None]).max(axis=0),
        "mean/N": _eps_rv_np(c, u, np.zeros_like(u), A, h, c_rv, mean_div=1),
        "c_rv twice": c_rv * h * h * (np.abs(A) / ninf[:, None]).max(axis=0),
    }
    for name, w in wrong.items():
        assert rel_inf(w, expect) > 1e-3, name


def _rv_case(c_uw):
    """Euler on 16 x 16 nodes, Dirichlet density-wave inflow and
    do-nothing outflow, a seeded non-smooth state and approx_du."""
    c = cases.build(configs.c2(), side=16, use_gpu=False, visc="residual", c_rv=0.8, c_uw=c_uw)
    N = c.N
    rng = np.random.default_rng(11)
    r = rng.uniform(0.0, 1.0, (4, N))
    u = np.ascontiguousarray(np.stack([1.0 + 0.2 * r[0], -3.0 + 0.5 * r[1], -0.2 + 0.1 * r[2], 6.5 + 0.2 * r[3]]))
    inflow = c.bc[0][1]
    # the first boundary pass sets the inflow nodes to the density wave at t = 0
    xb, yb = c.x[inflow], c.y[inflow]
    rho_b = 1.0 + 0.2 * np.sin(2.0 * np.pi * (xb + yb))
    u_seen = u.copy()
    u_seen[:, inflow] = np.stack([rho_b, rho_b, rho_b, 1.0 / 0.4 + rho_b])
    f1, f2, v1, v2, p = _euler_flux_np(u_seen)
    du_seen = -(np.einsum("is,vis->vi", c.wx, f1[:, c.idx]) + np.einsum("is,vis->vi", c.wy, f2[:, c.idx]))
    A = du_seen + rng.uniform(0.5, 1.5, (4, N)) * rng.choice([-1.0, 1.0], (4, N))
    eps_uw = c_uw * 0.5 * c.dx_avg * (np.sqrt(v1 * v1 + v2 * v2) + np.sqrt(1.4 * p / u_seen[0]))
    return c, u, u_seen, du_seen, A, eps_uw, inflow


def test_residual_viscosity_known_answer_with_boundary():
    """eps_rv on a non-smooth state with Dirichlet inflow nodes: the mean and
    the lexicographic max over the state after the first boundary pass, the
    residual against du before the second pass (the flux divergence, not the
    zeroed Dirichlet rows).  eps_uw is made large so that eps = eps_rv."""
    c, u, u_seen, du_seen, A, _, inflow = _rv_case(1e6)
    _, eps, code = c.prob.rhs(u.copy(), 0.0, success_iter=1, approx_du=A)
    h = c.dx_avg
    expect = _eps_rv_np(c, u_seen, du_seen, A, h, 0.8)
    assert np.all(code == 0)
    assert rel_inf(eps, expect) < 1e-10
    du_post = du_seen.copy()
    du_post[:, inflow] = 0.0
    f1, f2, _, _, _ = _euler_flux_np(u)
    du_pre = -(np.einsum("is,vis->vi", c.wx, f1[:, c.idx]) + np.einsum("is,vis->vi", c.wy, f2[:, c.idx]))
    dev = np.abs(u_seen - (u_seen.sum(1) / (4 * c.N))[:, None])
    assert np.argmax(dev[0]) != np.argmax(dev[1])  # the lexicographic and component-wise maxima differ
    wrong = {
        "h": expect / h,
        "dx_min": expect * (c.dx_min / h) ** 2,
        "mean/N": _eps_rv_np(c, u_seen, du_seen, A, h, 0.8, mean_div=1),
        "component-wise max": _eps_rv_np(c, u_seen, du_seen, A, h, 0.8, lex=False),
        "du after the second boundary pass": _eps_rv_np(c, u_seen, du_post, A, h, 0.8),
        "u before the first boundary pass": _eps_rv_np(c, u, du_pre, A, h, 0.8),
    }
    for name, w in wrong.items():
        assert rel_inf(w, expect) > 1e-4, name


def test_residual_viscosity_branches_known_answer():
    """eps = min(eps_rv, eps_uw) with code 0 where eps_rv < eps_uw, 1 else
    (update_visc!, hyperviscosity.jl:331-349), with c_uw chosen so that both
    branches occur; success_iter = 0 gives eps_uw with code 1 everywhere."""
    c, u, u_seen, du_seen, A, _, _ = _rv_case(1.0)
    rv = _eps_rv_np(c, u_seen, du_seen, A, c.dx_avg, 0.8)
    _, _, _, _, _, uw1, _ = _rv_case(1.0)
    c_uw = float(np.median(rv / uw1))
    c, u, u_seen, du_seen, A, uw, _ = _rv_case(c_uw)
    _, eps, code = c.prob.rhs(u.copy(), 0.0, success_iter=4, approx_du=A)
    expect_code = np.where(rv < uw, 0, 1)
    clear = np.abs(rv / uw - 1.0) > 1e-9
    assert 0.2 < np.mean(expect_code) < 0.8
    assert np.array_equal(code[clear], expect_code[clear])
    assert rel_inf(eps, np.minimum(rv, uw)) < 1e-10
    _, eps0, code0 = c.prob.rhs(u.copy(), 0.0, success_iter=0, approx_du=A)
    assert np.all(code0 == 1) and rel_inf(eps0, uw) < 1e-13


def test_host_viscosity_spacing_is_dx_avg():
    """The host API passes PointData's dx_avg (the mean nearest-neighbour
    distance) as h of both viscosity sources and dx_min to the
    hyperviscosity sources (hyperviscosity.jl:47, :116, :280, :324)."""
    from mfrbf.domain import PointData
    from mfrbf.sources import (SourceHyperviscosityFlyer, SourceHyperviscosityTominec,
                               SourceResidualViscosityTominec, SourceUpwindViscosityTominec)

    class Dom:
        pd = PointData(np.zeros((2, 2)), np.zeros((2, 1), np.int32), 0.01, 0.03)

    assert SourceResidualViscosityTominec(None, None, Dom).h == 0.03
    assert SourceUpwindViscosityTominec(None, None, Dom).h == 0.03
    assert SourceHyperviscosityFlyer(None, None, Dom, k=2, c=3.0).gamma == 3.0 * 0.01 ** 4
    assert SourceHyperviscosityTominec(None, None, Dom, c=2.0).gamma == 2.0 * 0.01 ** 4.5


# ----------------------------------------------------------------- history
def test_history_closed_form():
    """test/history_test.jl:1-32: linear data gives derivative 1 (order 5)."""
    P = 6
    th = np.zeros(P)
    sh = np.zeros(P * P)
    ad = np.zeros(P)
    u = np.zeros(P)
    t = 0.0
    for i in range(P):
        t += 1.0
        u += 1.0
        O.history_update(5, 0, t, u.copy(), th, sh, ad)
    O.history_update(5, P, t + 1.0, u + 1.0, th, sh, ad)
    assert np.allclose(ad, 1.0, rtol=1e-12)


def test_time_derivative_weights_bdf():
    dt = 0.01
    t = 1.0 - dt * np.arange(3)
    w = O.time_deriv_weights(t)
    assert np.allclose(w, np.array([1.5, -2.0, 0.5]) / dt, rtol=1e-9)
    w2 = O.time_deriv_weights(t[:2])
    assert np.allclose(w2, np.array([1.0, -1.0]) / dt, rtol=1e-12)


# ------------------------------------------------------------------ SSPRK3
def test_ssprk33_is_third_order_taylor_for_linear_systems():
    spec = configs.c1()
    c = cases.build(spec, side=24, use_gpu=False)
    N = c.N
    prob = O.Problem(c.x, c.y, c.ptr, c.idx, c.wx, c.wy, O.EQ_ADVECTION, 1, adv=(1.0, 1.0))
    Dx = sp.csr_matrix((c.wx.ravel(), c.idx.ravel(), np.arange(N + 1) * c.k), shape=(N, N))
    Dy = sp.csr_matrix((c.wy.ravel(), c.idx.ravel(), np.arange(N + 1) * c.k), shape=(N, N))
    Lop = -(Dx + Dy)
    u0 = c.u0.copy()
    dt = 0.3 * c.dt
    u = u0.copy()
    prob.ssprk33(u, 0.0, dt, 1)
    v = u0[0]
    t1 = dt * (Lop @ v)
    t2 = dt * (Lop @ t1) / 2
    t3 = dt * (Lop @ t2) / 3
    assert rel_inf(u[0], v + t1 + t2 + t3) < 1e-13


def test_ssprk33_history_interleave_hand_stages():
    """Three SSPRK33 steps with the HistoryCallback (approx_order 3) and the
    residual viscosity, stage by stage from single rhs! calls: the callback
    runs at init (success_iter 0: approx_du = 0) and after each step with
    success_iter 1, 2, 3; approx_du is then the 2-, 3- and 4-point backward
    difference (1, -1)/dt, (3/2, -2, 1/2)/dt, (11/6, -3, 3/2, -1/3)/dt
    (history.jl:91-152); the stages of step n see the success_iter and
    approx_du of the callback after step n-1; the FSAL evaluation at the
    end of a step runs before that step's callback and is not redone after
    it (u_modified!(integrator, false), history.jl:73)."""
    c = cases.build(configs.c2(), side=14, use_gpu=False, visc="residual")
    u0 = c.u0.copy()
    dt = c.dt
    u = u0.copy()
    c.prob.ssprk33(u, 0.0, dt, 3, approx_order=3)
    got_u, got_ad = u, c.prob.approx_du.reshape(u0.shape).copy()

    bdf = {2: [1.0, -1.0], 3: [1.5, -2.0, 0.5], 4: [11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0]}

    def run(redo_fsal):
        f = lambda v, t, si, ad: c.prob.rhs(v, t, success_iter=si, approx_du=ad)[0]  # noqa: E731
        hist = []  # most recent first
        un = u0.copy()
        ad = np.zeros_like(un)
        hist.insert(0, un.copy())
        k = f(un, 0.0, 0, ad)
        t = 0.0
        for n in range(3):
            si = n
            a = un + dt * k
            k = f(a, t + dt, si, ad)
            b = (3.0 * un + (a + dt * k)) / 4.0
            k = f(b, t + dt / 2.0, si, ad)
            un = (un + 2.0 * b + 2.0 * dt * k) / 3.0
            k = f(un, t + dt, si, ad)  # FSAL, before the callback
            t = t + dt
            hist.insert(0, un.copy())
            q = min(n + 2, 4)
            ad = sum(w * h for w, h in zip(bdf[q], hist[:q])) / dt
            if redo_fsal:
                k = f(un, t, n + 1, ad)
        return un, ad

    u_ref, ad_ref = run(False)
    assert rel_inf(got_u, u_ref) < 1e-13
    assert rel_inf(got_ad, ad_ref) < 1e-8  # backward differences cancel about 4 digits
    u_bad, _ = run(True)
    assert rel_inf(u_bad, u_ref) > 1e-9  # the interleave is visible in the state


# ----------------------------------------------------------------- limiter
def test_limiter_hand_example():
    # 3 nodes, stencils of 2: node 0 -> (0, 1), node 1 -> (1, 0), node 2 -> (2, 1)
    u = np.array([[0.1, 1.0, 2.0]])
    ptr = np.array([0, 2, 4, 6])
    idx = np.array([0, 1, 1, 0, 2, 1])
    O.limiter_zhang_shu(u, ptr, idx, 0.5, O.VAR_DENSITY)
    # node 0 and 1 see min 0.1 < 0.5, mean 0.55, theta = (0.55-0.5)/(0.55-0.1)
    th = (0.55 - 0.5) / (0.55 - 0.1)
    assert np.allclose(u[0, :2], [th * 0.1 + (1 - th) * 0.55, th * 1.0 + (1 - th) * 0.55], rtol=1e-15)
    assert u[0, 2] == 2.0
    assert u[0].min() >= 0.5 - 1e-15 or u[0, 0] >= 0.5 - 1e-12


# ----------------------------------------------------------------- adjoint
def test_vjp_matches_finite_differences():
    c = cases.build(configs.c2(), side=20, use_gpu=False)
    u0 = c.u0
    rng = np.random.default_rng(5)
    lam = rng.standard_normal(u0.shape)
    v = rng.standard_normal(u0.shape)
    ub = c.prob.vjp(u0, lam, 0.0)
    eps = 1e-6
    fp, _, _ = c.prob.rhs(np.ascontiguousarray(u0 + eps * v), 0.0)
    fm, _, _ = c.prob.rhs(np.ascontiguousarray(u0 - eps * v), 0.0)
    jv = (fp - fm) / (2 * eps)
    lhs, rhs = float((lam * jv).sum()), float((ub * v).sum())
    assert abs(lhs - rhs) < 1e-6 * abs(rhs)


@pytest.mark.parametrize("bc", [("nothing", "nothing"), ("dirichlet", "nothing")])
def test_vjp_matches_forward_mode(bc):
    """Dot-product test <lam, J v> = <J' lam, v> at rounding level: the
    oracle's VJP against the forward-mode J v above.  A Dirichlet node's
    state is overwritten by the boundary value (so it does not depend on u)
    and its du is zeroed, so v and J v vanish there."""
    c = cases.build(configs.c2(), side=20, use_gpu=False, bc=bc)
    rng = np.random.default_rng(7)
    lam = rng.standard_normal(c.u0.shape)
    v = rng.standard_normal(c.u0.shape)
    ub = c.prob.vjp(c.u0.copy(), lam, 0.0)
    ubc = c.u0.copy()
    c.prob.rhs(ubc, 0.0)  # applies the first boundary pass to ubc in place
    fixed = c.bc[0][1] if bc[0] == "dirichlet" else np.zeros(0, np.int64)
    vv = v.copy()
    vv[:, fixed] = 0.0
    jv = cases.euler_jvp(c, ubc, vv)
    jv[:, fixed] = 0.0
    lhs, rhs = float((lam * jv).sum()), float((ub * v).sum())
    scale = float(np.abs(lam * jv).sum())
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)
    # and node by node: the VJP is the transpose of the same linear map
    e = np.zeros_like(v)
    q = int(rng.integers(c.N))
    e[1, q] = 1.0
    ee = e.copy()
    ee[:, fixed] = 0.0
    jq = cases.euler_jvp(c, ubc, ee)
    jq[:, fixed] = 0.0
    assert abs(float((lam * jq).sum()) - ub[1, q]) <= 1e-12 * float(np.abs(lam * jq).sum() + 1e-300)


@pytest.mark.parametrize("visc", ["residual", "upwind"])
def test_rhs_par_matches_serial(visc):
    """The CPU baseline's OpenMP rhs! (row gathers and column dots split over
    threads) against the serial oracle: the same du and the same boundary
    mutation of u, to rounding (its sums run in another order)."""
    c = cases.build(configs.c2(), side=40, use_gpu=False, visc=visc)
    u1 = c.u0 * 1.0001
    th, sh, ad = np.zeros(4), np.zeros(4 * c.u0.size), np.zeros(c.u0.size)
    O.history_update(3, 0, 0.0, c.u0.ravel().copy(), th, sh, ad)
    O.history_update(3, 1, 0.001, u1.ravel().copy(), th, sh, ad)
    c.prob.approx_du[:] = ad
    for si in (0, 1):  # 0: the upwind eps only; 1: the residual eps where finite
        c.prob.cfg.success_iter = si
        us, up = c.u0.copy(), c.u0.copy()
        ds, _, _ = c.prob.rhs(us, 0.001)
        dp = c.prob.rhs_par(up, 0.001, nthreads=4)
        assert np.array_equal(us, up)  # the first boundary pass
        assert rel_inf(dp, ds) < 1e-12
        assert np.abs(ds).max() > 0.0
    hv = cases.build(configs.c2(), side=16, use_gpu=False, hv=(2, 1.0))
    with pytest.raises(ValueError):
        hv.prob.rhs_par(hv.u0.copy(), 0.0)


def test_tominec_hyperviscosity_matches_explicit_product():
    """SourceHyperviscosityTominec (hyperviscosity.jl:80-134): the reference
    builds H = lap' * lap explicitly (lap = Dxx + Dyy) and adds -gamma H u.
    The oracle applies lap' (lap u); pinned here against scipy's explicit
    product, and lap against exact second derivatives of quadratics."""
    spec = configs.c2()
    x, y, _, _, _ = O.cloud_square(24, spec.seed)
    N, k = len(x), 30
    idx, _ = O.knn_grid(x, y, k)
    ptr = O.ell_ptr(N, k)
    hx, hy, rc = O.weights(x, y, ptr, idx, 5, 3, 2)
    assert rc == 0
    lap = (hx + hy).reshape(N, k)
    wx, wy, rc = O.weights(x, y, ptr, idx, 5, 3, 1)
    assert rc == 0
    Lm = sp.csr_matrix((lap.ravel(), idx.ravel(), np.arange(N + 1) * k), shape=(N, N))
    # lap is exact on quadratics: lap(x^2 + 3xy - y^2) = 2 - 2 = 0, lap(x^2 + y^2) = 4
    assert np.abs(Lm @ (x * x + y * y) - 4.0).max() < 1e-6
    assert np.abs(Lm @ (x * x + 3 * x * y - y * y)).max() < 1e-6
    g = 1e-4 * (1.0 / 23) ** 4.5
    u0 = np.stack([1.0 + 0.2 * np.sin(2 * np.pi * (x + y)), 1.0 + 0 * x, 1.0 + 0 * x, 2.5 + x * y])
    plain = O.Problem(x, y, ptr, idx, wx, wy, O.EQ_EULER, 4)
    tom = O.Problem(x, y, ptr, idx, wx, wy, O.EQ_EULER, 4, lap=lap, tom_gamma=g)
    du0, _, _ = plain.rhs(u0.copy(), 0.0)
    du1, _, _ = tom.rhs(u0.copy(), 0.0)
    H = (Lm.T @ Lm).tocsr()
    ref = -g * (H @ u0.T).T
    assert rel_inf(du1 - du0, ref) < 1e-12
    # a constant state feels no hyperviscosity
    c = np.ones_like(u0)
    dc1, _, _ = tom.rhs(c.copy(), 0.0)
    dc0, _, _ = plain.rhs(c.copy(), 0.0)
    assert np.abs(dc1 - dc0).max() < 1e-12 * np.abs(g * (abs(H) @ np.ones(N))).max() + 1e-18


def test_source_order_flyer_tominec_upwind():
    """SourceTerms applies its sources in keyword order (generic_sources.jl:
    7-18).  With the upwind viscosity (eps from u alone) every order gives
    the same du up to rounding, and each source adds its own term: the
    rhs! with all three equals the flux-only rhs! plus the three
    single-source differences."""
    spec = configs.c2()
    x, y, _, _, _ = O.cloud_square(20, spec.seed)
    N, k = len(x), 30
    idx, d2 = O.knn_grid(x, y, k)
    ptr = O.ell_ptr(N, k)
    dx_min, dx_avg = O.dx_stats(d2[:, 1])
    wx, wy, rc = O.weights(x, y, ptr, idx, 5, 3, 1)
    hx, hy, _ = O.weights(x, y, ptr, idx, 5, 3, 4)
    lx, ly, _ = O.weights(x, y, ptr, idx, 5, 3, 2)
    H, lap = (hx + hy).reshape(N, k), (lx + ly).reshape(N, k)
    gf, gt = dx_min ** 4, dx_min ** 4.5
    u0 = np.stack([1.0 + 0.2 * np.sin(2 * np.pi * (x + y)), 0.3 + 0 * x, -0.2 + 0.1 * y, 2.5 + x * y])

    def du(order, fly=True, tom=True, rv=True):
        p = O.Problem(x, y, ptr, idx, wx, wy, O.EQ_EULER, 4, hv=H if fly else None, hv_gamma=gf,
                      rv_mode=O.RV_UPWIND if rv else O.RV_NONE, dx_avg=dx_avg,
                      lap=lap if tom else None, tom_gamma=gt, order=order)
        return p.rhs(u0.copy(), 0.0)[0]

    F, T, R = O.SRC_FLYER, O.SRC_TOMINEC, O.SRC_VISCOSITY
    ref = du([F, T, R])
    scale = np.abs(ref).max()
    for order in ([F, R, T], [T, F, R], [T, R, F], [R, F, T], [R, T, F]):
        assert np.abs(du(order) - ref).max() < 1e-12 * scale
    base = du([F], fly=False, tom=False, rv=False)
    parts = (du([F], tom=False, rv=False) - base) + (du([T], fly=False, rv=False) - base) + \
        (du([R], fly=False, tom=False) - base)
    assert np.abs(base + parts - ref).max() < 1e-12 * scale
    # and the Tominec part is not negligible here
    assert np.abs(du([T], fly=False, rv=False) - base).max() > 1e-6 * scale


def _igr_case(side=24):
    spec = configs.c2()
    x, y, _, _, _ = O.cloud_square(side, spec.seed)
    N, k = len(x), 30
    idx, d2 = O.knn_grid(x, y, k)
    ptr = O.ell_ptr(N, k)
    wx, wy, rc = O.weights(x, y, ptr, idx, 5, 3, 1)
    assert rc == 0
    Dx = sp.csr_matrix((wx, idx.ravel(), np.arange(N + 1) * k), shape=(N, N))
    Dy = sp.csr_matrix((wy, idx.ravel(), np.arange(N + 1) * k), shape=(N, N))
    dx_min, _ = O.dx_stats(d2[:, 1])
    return x, y, ptr, idx, wx, wy, Dx, Dy, dx_min


def test_igr_rhs_closed_form():
    """update_igr_rhs! (IGR.jl:117-157) on a state with constant density and
    a linear velocity field: the derivative weights are exact on linear
    functions, so trace = D_x rho + D_y v1 = b and trace_sq = a^2 + 2 b c + d^2
    for v = (a x + b y, c x + d y) at every node."""
    x, y, ptr, idx, wx, wy, Dx, Dy, _ = _igr_case()
    a_, b_, c_, d_ = 0.3, -0.7, 1.1, 0.45
    rho = 1.3 + 0 * x
    u = np.stack([rho, rho * (a_ * x + b_ * y), rho * (c_ * x + d_ * y), 2.5 + 0 * x])
    alpha = 0.25
    p = O.Problem(x, y, ptr, idx, wx, wy, O.EQ_EULER, 4, igr_alpha=alpha)
    b, rinv = p.igr_rhs(u)
    expect = alpha * (b_ * b_ + a_ * a_ + 2 * b_ * c_ + d_ * d_)
    assert np.abs(b - expect).max() < 1e-8 * abs(expect)
    assert np.array_equal(rinv, 1.0 / rho)


def test_igr_operator_and_cg():
    """The IGRCompositeMatrix product (IGR.jl:50-66) against the explicit
    sparse matrix diag(1/rho) - alpha (Dx diag(1/rho) Dx + Dy diag(1/rho) Dy),
    and cg! against a direct solve where CG converges (alpha = dx^2; the
    reference's default alpha = 1 makes the operator indefinite and CG does
    not converge), with the early exit for b = 0."""
    x, y, ptr, idx, wx, wy, Dx, Dy, dx_min = _igr_case()
    N = len(x)
    rng = np.random.default_rng(4)
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * (x + y))
    u = np.stack([rho, rho * (1.0 + 0.1 * x), rho * (1.0 - 0.2 * y * y), 2.5 + rho])
    alpha = dx_min ** 2
    p = O.Problem(x, y, ptr, idx, wx, wy, O.EQ_EULER, 4, igr_alpha=alpha)
    b, rinv = p.igr_rhs(u)
    R = sp.diags(rinv)
    A = (R - alpha * (Dx @ R @ Dx + Dy @ R @ Dy)).toarray()
    v = rng.standard_normal(N)
    assert rel_inf(p.igr_apply(rinv, v), A @ v) < 1e-12
    xs = np.linalg.solve(A, b)
    x60, it = p.igr_cg(rinv, b, maxiter=60)
    assert it == 60 or np.linalg.norm(b - A @ x60) <= np.sqrt(np.finfo(float).eps) * np.linalg.norm(b) * 1.01
    assert rel_inf(x60, xs) < 1e-5
    x5, it5 = p.igr_cg(rinv, b, maxiter=5)
    assert it5 == 5 and rel_inf(x5, xs) > rel_inf(x60, xs)
    z, it0 = p.igr_cg(rinv, 0 * b, maxiter=20)
    assert it0 == 0 and np.all(z == 0.0)


def test_igr_source_terms_in_du():
    """The source adds -Dx sigma to du_m1 and -Dy sigma to du_m2 only
    (flux_igr, IGR.jl:193-239); sigma is the CG solution from 0."""
    x, y, ptr, idx, wx, wy, Dx, Dy, dx_min = _igr_case()
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * (x + y))
    u = np.stack([rho, rho, rho, 2.5 + rho])
    alpha = 20 * dx_min ** 2
    plain = O.Problem(x, y, ptr, idx, wx, wy, O.EQ_EULER, 4)
    igr = O.Problem(x, y, ptr, idx, wx, wy, O.EQ_EULER, 4, igr_alpha=alpha, igr_maxiter=20)
    du0, _, _ = plain.rhs(u.copy(), 0.0)
    du1, _, _ = igr.rhs(u.copy(), 0.0)
    b, rinv = igr.igr_rhs(u)
    sig, _ = igr.igr_cg(rinv, b, maxiter=20)
    assert np.array_equal(igr.igr_sigma, sig)
    d = du1 - du0
    assert np.all(d[0] == 0.0) and np.all(d[3] == 0.0)
    assert np.abs(d[1] + Dx @ sig).max() < 1e-12 * np.abs(Dx @ sig).max()
    assert np.abs(d[2] + Dy @ sig).max() < 1e-12 * np.abs(Dy @ sig).max()


@pytest.mark.parametrize("fixture", ["c1_advection.npz", "euler_rv_1024.npz"])
def test_fixture_reproduced(fixture):
    """The oracle still computes the committed fixtures bit for bit: C1
    (cloud, stencils, weights, du at t = 0, the state after 200 SSPRK33
    steps) and Euler with the residual viscosity (3 steps with the history
    callback; eps, codes and approx_du of the last rhs!)."""
    sys.path.insert(0, os.path.join(GOLDEN))
    import make_fixtures

    ref = np.load(os.path.join(GOLDEN, fixture))
    now = make_fixtures.FIXTURES[fixture]()
    assert sorted(ref.files) == sorted(now)
    for name in ref.files:
        assert np.array_equal(ref[name], now[name]), name
    assert np.all(np.isfinite(ref["u_final"]))
    if fixture == "euler_rv_1024.npz":
        assert np.any(ref["eps_code"] == 0)  # eps_rv chosen somewhere: the residual path is live

