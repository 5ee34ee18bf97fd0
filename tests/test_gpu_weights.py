"""RBF-FD weights from the null-space solver on the matrix cores (through the
C ABI) against the oracle's Bunch-Kaufman restatement of the reference's
M \\ rhs, and against a 40-digit solve of the same saddle systems.

The two factorizations round differently, so the bar is not bit equality:
per stencil max|w - w_oracle| <= 1e-10 max|w_oracle| (the saddle systems'
condition numbers reach about 1e9, and the Bunch-Kaufman solution itself is
off by up to 4e-11 of that scale against the 40-digit one); the weights
reproduce the derivatives of every monomial of the basis to 1e-9, as the
oracle's do.  MF_TUNE_WEIGHTS_SOLVER 1 (Bunch-Kaufman on the device) still
gives the oracle's weights."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_inf
from mfrbf import _lib as L
from mfrbf import configs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _per_stencil_err(w, wo, ptr):
    e = 0.0
    for a, b in zip(ptr[:-1], ptr[1:]):
        e = max(e, float(np.abs(w[a:b] - wo[a:b]).max() / np.abs(wo[a:b]).max()))
    return e


@pytest.mark.parametrize("p,deg,k,kd", [(3, 2, 15, 1), (5, 3, 30, 1), (5, 3, 30, 4), (3, 3, 20, 1),
                                        (5, 3, 30, 2), (5, 3, 50, 1), (7, 4, 40, 1),
                                        (7, 3, 30, 1), (3, 3, 30, 2), (5, 2, 15, 4),
                                        (3, 3, 40, 1), (5, 3, 44, 1), (3, 3, 50, 2)])
def test_ns_operator_vs_oracle(mf, p, deg, k, kd):
    """mf_compute_operator with the null-space solver, every basis the
    configurations and the hyperviscosity sources use (and a degree-4 one):
    per-stencil agreement with the oracle and polynomial reproduction.  The
    kernels compiled for 15 nodes of degree 2 and 30 of degree 3 (PHS powers
    3, 5, 7), and the 33..64-node variants with the 10 monomials of degree 3
    compiled in, are among them."""
    _say("null-space weights p=%d deg=%d k=%d kd=%d" % (p, deg, k, kd))
    x, y, _, _, _ = O.cloud_square(40, 20261202)
    N = len(x)
    idx, _ = O.knn_grid(x, y, k)
    ctx = L.Context(0)
    ctx.set_tuning(L.MF_TUNE_WEIGHTS_SOLVER, 2)
    ctx.set_points(x, y)
    ctx.set_stencil(idx)
    wx, wy = ctx.compute_operator(p, deg, kd, L.MF_OP_FLUX, want=True, k=k)
    ox, oy, rc = O.weights(x, y, O.ell_ptr(N, k), idx, p, deg, kd)
    assert rc == 0
    ptr = O.ell_ptr(N, k)
    ex, ey = _per_stencil_err(wx.ravel(), ox, ptr), _per_stencil_err(wy.ravel(), oy, ptr)
    _say("  per-stencil rel diff %.2e %.2e" % (ex, ey))
    assert ex < 1e-10 and ey < 1e-10
    tol_x = 1e-9 + 1e-10 * np.abs(wx).max()
    tol_y = 1e-9 + 1e-10 * np.abs(wy).max()
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            f = x ** a * y ** b
            # d^kd/dx^kd x^a y^b = a (a-1) ... (a-kd+1) x^(a-kd) y^b
            dfx = np.prod(np.arange(a - kd + 1, a + 1)) * x ** (a - kd) * y ** b if a >= kd else 0 * x
            dfy = np.prod(np.arange(b - kd + 1, b + 1)) * x ** a * y ** (b - kd) if b >= kd else 0 * x
            assert np.abs((wx * f[idx]).sum(1) - dfx).max() < tol_x
            assert np.abs((wy * f[idx]).sum(1) - dfy).max() < tol_y


def test_ns_basis_below_cpd_order_uses_bunch_kaufman(mf):
    """r^7 with degree-2 monomials is below the order (4) of conditional
    positive definiteness, where the null-space K need not be positive
    definite: the null-space setting then solves by Bunch-Kaufman, and the
    weights equal those of MF_TUNE_WEIGHTS_SOLVER 1."""
    _say("null-space setting below the CPD order")
    x, y, _, _, _ = O.cloud_square(40, 20261202)
    idx, _ = O.knn_grid(x, y, 15)
    out = []
    for solver in (2, 1):
        ctx = L.Context(0)
        ctx.set_tuning(L.MF_TUNE_WEIGHTS_SOLVER, solver)
        ctx.set_points(x, y)
        ctx.set_stencil(idx)
        out.append(ctx.compute_operator(7, 2, 2, L.MF_OP_FLUX, want=True, k=15))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_ns_small_stencils_use_bunch_kaufman(mf):
    """A stencil of fewer nodes than twice the monomial count (15 nodes with
    the 10 of degree 3) is solved by Bunch-Kaufman under the null-space
    setting: the operator equals that of MF_TUNE_WEIGHTS_SOLVER 1 bit for
    bit."""
    _say("null-space setting, 15-node stencils of degree 3")
    x, y, _, _, _ = O.cloud_square(40, 20261202)
    idx, _ = O.knn_grid(x, y, 15)
    out = []
    for solver in (2, 1):
        ctx = L.Context(0)
        ctx.set_tuning(L.MF_TUNE_WEIGHTS_SOLVER, solver)
        ctx.set_points(x, y)
        ctx.set_stencil(idx)
        out.append(ctx.compute_operator(5, 3, 1, L.MF_OP_FLUX, want=True, k=15))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_ns_ragged_all_below_null_space_size(mf):
    """A ragged batch whose every stencil is below twice the monomial count
    (12 to 19 nodes for 10 monomials) runs only the Bunch-Kaufman launches:
    its weights equal those of MF_TUNE_WEIGHTS_SOLVER 1 bit for bit."""
    _say("ragged weights, every stencil below the null-space size")
    x, y, _, _, _ = O.cloud_square(32, 20261211)
    N = len(x)
    sizes = 12 + np.arange(N) % 8
    big, _ = O.knn_grid(x, y, 19)
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    idx = np.concatenate([big[e, :sizes[e]] for e in range(N)]).astype(np.int32)
    out = []
    for solver in (0, 1):
        ctx = L.Context(0)
        ctx.set_tuning(L.MF_TUNE_WEIGHTS_SOLVER, solver)
        ctx.set_points(x, y)
        ctx.weights_ragged_set(ptr, idx)
        ctx.weights_ragged_run(5, 3, 1)
        out.append(ctx.weights_ragged_get())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_ns_ragged_c3_sample_vs_oracle(mf):
    """C3's stencil mix (n = 30, 8 % with n in 40..50; here 20 % so that
    every size occurs often) on a 96 x 96 cloud, all sizes in one launch:
    per stencil within 1e-10 of the oracle; the Bunch-Kaufman device path
    (MF_TUNE_WEIGHTS_SOLVER 1) reproduces the oracle's bits."""
    _say("null-space ragged weights, C3 mix")
    x, y, _, _, _ = O.cloud_square(96, 20261204)
    N = len(x)
    sizes = configs.ragged_sizes(N, 20261204, p_boundary=0.2)
    big, _ = O.knn_grid(x, y, 50)
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    idx = np.concatenate([big[e, :sizes[e]] for e in range(N)]).astype(np.int32)
    ox, oy, rc = O.weights(x, y, ptr, idx, 5, 3, 1)
    assert rc == 0
    ctx = L.Context(0)
    ctx.set_points(x, y)
    ctx.weights_ragged_set(ptr, idx)
    ctx.weights_ragged_run(5, 3, 1)
    wx, wy = ctx.weights_ragged_get()
    ex, ey = _per_stencil_err(wx, ox, ptr), _per_stencil_err(wy, oy, ptr)
    _say("  per-stencil rel diff %.2e %.2e, norm-wise %.2e" % (ex, ey, rel_inf(wx, ox)))
    assert ex < 1e-10 and ey < 1e-10
    ctx.set_tuning(L.MF_TUNE_WEIGHTS_SOLVER, 1)
    ctx.weights_ragged_run(5, 3, 1)
    bx, by = ctx.weights_ragged_get()
    assert rel_inf(bx, ox) < 1e-12 and rel_inf(by, oy) < 1e-12


def test_ns_ragged_extreme_sizes(mf):
    """Ragged stencils at the edges of the supported range and of each
    path, interleaved: 11 nodes (one more than the 10 monomials of degree
    3), 12 and 19 (below twice the monomial count: Bunch-Kaufman, since the
    Cholesky QR of a nearly square P loses digits, 3e-7 at 11 nodes), 20
    (the smallest null-space size), 31, 32 (the largest of the two-tile
    variant), 33, 48, 49 and 64 (the largest stencil): per stencil within
    1e-10 of the oracle.  (An LU solve of the same systems differs from the
    oracle's by up to 2.5e-11 per stencil, largest at 11 and 64 nodes.)"""
    _say("null-space ragged weights, extreme sizes")
    x, y, _, _, _ = O.cloud_square(40, 20261208)
    N = len(x)
    choices = np.array([11, 12, 19, 20, 31, 32, 33, 48, 49, 64])
    sizes = choices[np.arange(N) % len(choices)]
    big, _ = O.knn_grid(x, y, 64)
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    idx = np.concatenate([big[e, :sizes[e]] for e in range(N)]).astype(np.int32)
    ox, oy, rc = O.weights(x, y, ptr, idx, 5, 3, 1)
    assert rc == 0
    ctx = L.Context(0)
    ctx.set_points(x, y)
    ctx.weights_ragged_set(ptr, idx)
    ctx.weights_ragged_run(5, 3, 1)
    wx, wy = ctx.weights_ragged_get()
    for n in choices:
        sel = np.nonzero(sizes == n)[0]
        sp = np.concatenate([[0], np.cumsum(sizes[sel])])
        gx = np.concatenate([wx[ptr[e]:ptr[e + 1]] for e in sel])
        gox = np.concatenate([ox[ptr[e]:ptr[e + 1]] for e in sel])
        gy = np.concatenate([wy[ptr[e]:ptr[e + 1]] for e in sel])
        goy = np.concatenate([oy[ptr[e]:ptr[e + 1]] for e in sel])
        ex, ey = _per_stencil_err(gx, gox, sp), _per_stencil_err(gy, goy, sp)
        _say("  n=%d: per-stencil rel diff %.2e %.2e" % (n, ex, ey))
        assert ex < 1e-10 and ey < 1e-10, n


def test_ns_closer_to_exact_than_tolerance(mf):
    """On a few C3 stencils (interior n = 30, boundary n = 40..50, and
    one-sided stencils at the domain edge) the null-space weights are within
    1e-11 of a 40-digit solve of the saddle system in the scaled
    coordinates, as close as the oracle's Bunch-Kaufman weights are."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    side = 64
    x, y, _, _, _ = O.cloud_square(side, configs.SEED0 + 3)
    N = len(x)
    big, _ = O.knn_grid(x, y, 50)
    picks = [(side * side // 2 + side // 2, 30), (side * 10 + 20, 45), (3, 30), (side * 7, 40), (N - 1, 50)]
    ptr = np.concatenate([[0], np.cumsum([n for _, n in picks])]).astype(np.int64)
    idx = np.concatenate([big[e, :n] for e, n in picks]).astype(np.int32)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    ctx.weights_ragged_set(ptr, idx)
    ctx.weights_ragged_run(5, 3, 1)
    wx, _ = ctx.weights_ragged_get()
    ox, _, _ = O.weights(x, y, ptr, idx, 5, 3, 1)
    mons = [(d - t, t) for d in range(4) for t in range(d + 1)]
    for s, (e, n) in enumerate(picks):
        nodes = idx[ptr[s]:ptr[s + 1]]
        X = [(mp.mpf(x[q]) - mp.mpf(x[nodes[0]]), mp.mpf(y[q]) - mp.mpf(y[nodes[0]])) for q in nodes]
        sx = 1 / max(abs(v[0]) for v in X)
        sy = 1 / max(abs(v[1]) for v in X)
        Xs = [(v[0] * sx, v[1] * sy) for v in X]
        m = n + len(mons)
        M = mp.zeros(m, m)
        rhs = mp.zeros(m, 1)
        for i in range(n):
            for j in range(n):
                M[i, j] = mp.sqrt((Xs[i][0] - Xs[j][0]) ** 2 + (Xs[i][1] - Xs[j][1]) ** 2) ** 5
            for t, (ax, ay) in enumerate(mons):
                M[i, n + t] = M[n + t, i] = Xs[i][0] ** ax * Xs[i][1] ** ay
            px, py = (-Xs[i][0], -Xs[i][1]) if i else (mp.mpf(2) ** -52, mp.mpf(2) ** -52)
            rhs[i] = 5 * px * mp.sqrt(px * px + py * py) ** 3
        rhs[n + 1] = 1  # d/dx of the monomial x at the centre
        w = mp.lu_solve(M, rhs)
        exact = np.array([float(w[i] * sx) for i in range(n)])
        scale = np.abs(exact).max()
        e_ns = np.abs(wx[ptr[s]:ptr[s + 1]] - exact).max() / scale
        e_bk = np.abs(ox[ptr[s]:ptr[s + 1]] - exact).max() / scale
        _say("  stencil %d (n=%d): null-space %.2e, Bunch-Kaufman %.2e vs 40 digits" % (e, n, e_ns, e_bk))
        assert e_ns < 1e-11


def test_ns_r7_large_stencils_vs_40_digits(mf):
    """r^7 with degree 3 on stencils of 44 and 50 nodes (the larger
    variants with the 10 monomials compiled in), where the saddle systems
    are ill-conditioned enough that the null-space and Bunch-Kaufman weights
    differ by several 1e-10: against a 40-digit solve both are about
    equally close, and the null-space error stays below 1e-9."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    side = 64
    x, y, _, _, _ = O.cloud_square(side, configs.SEED0 + 3)
    N = len(x)
    big, _ = O.knn_grid(x, y, 50)
    picks = [(side * side // 2 + side // 2, 44), (side * 10 + 20, 50), (side * 7, 44), (N - 1, 50)]
    ptr = np.concatenate([[0], np.cumsum([n for _, n in picks])]).astype(np.int64)
    idx = np.concatenate([big[e, :n] for e, n in picks]).astype(np.int32)
    ctx = L.Context(0)
    ctx.set_points(x, y)
    ctx.weights_ragged_set(ptr, idx)
    ctx.weights_ragged_run(7, 3, 1)
    wx, _ = ctx.weights_ragged_get()
    ox, _, _ = O.weights(x, y, ptr, idx, 7, 3, 1)
    mons = [(d - t, t) for d in range(4) for t in range(d + 1)]
    for s, (e, n) in enumerate(picks):
        nodes = idx[ptr[s]:ptr[s + 1]]
        X = [(mp.mpf(x[q]) - mp.mpf(x[nodes[0]]), mp.mpf(y[q]) - mp.mpf(y[nodes[0]])) for q in nodes]
        sx = 1 / max(abs(v[0]) for v in X)
        sy = 1 / max(abs(v[1]) for v in X)
        Xs = [(v[0] * sx, v[1] * sy) for v in X]
        m = n + len(mons)
        M = mp.zeros(m, m)
        rhs = mp.zeros(m, 1)
        for i in range(n):
            for j in range(n):
                M[i, j] = mp.sqrt((Xs[i][0] - Xs[j][0]) ** 2 + (Xs[i][1] - Xs[j][1]) ** 2) ** 7
            for t, (ax, ay) in enumerate(mons):
                M[i, n + t] = M[n + t, i] = Xs[i][0] ** ax * Xs[i][1] ** ay
            px, py = (-Xs[i][0], -Xs[i][1]) if i else (mp.mpf(2) ** -52, mp.mpf(2) ** -52)
            rhs[i] = 7 * px * mp.sqrt(px * px + py * py) ** 5
        rhs[n + 1] = 1  # d/dx of the monomial x at the centre
        w = mp.lu_solve(M, rhs)
        exact = np.array([float(w[i] * sx) for i in range(n)])
        scale = np.abs(exact).max()
        e_ns = np.abs(wx[ptr[s]:ptr[s + 1]] - exact).max() / scale
        e_bk = np.abs(ox[ptr[s]:ptr[s + 1]] - exact).max() / scale
        _say("  stencil %d (n=%d): null-space %.2e, Bunch-Kaufman %.2e vs 40 digits" % (e, n, e_ns, e_bk))
        assert e_ns < 1e-9 and e_ns < 10 * e_bk + 1e-12
