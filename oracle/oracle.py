"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  The C sources
cite the reference (MeshfreeTrixi.jl) file:line they restate.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EQ_ADVECTION, EQ_EULER = 0, 1
BC_DIRICHLET, BC_SLIP, BC_NOTHING = 0, 1, 2
FN_CONST, FN_PULSE, FN_WAVE, FN_GRADIENT, FN_TABLE = 0, 1, 2, 3, 4
RV_NONE, RV_RESIDUAL, RV_UPWIND = 0, 1, 2
VAR_DENSITY, VAR_PRESSURE = 0, 1
SRC_FLYER, SRC_TOMINEC, SRC_VISCOSITY, SRC_IGR = 1, 2, 3, 4


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_i8p = C.POINTER(C.c_int8)
_ip = C.POINTER(C.c_int)


class RhsCfg(C.Structure):
    _fields_ = [
        ("N", C.c_int64), ("nvar", C.c_int), ("eq", C.c_int),
        ("adv_a", C.c_double * 2), ("gamma", C.c_double),
        ("px", _dp), ("py", _dp),
        ("colptr", _i64p), ("rowval", _i32p), ("dx", _dp), ("dy", _dp),
        ("hv", _dp), ("hv_gamma", C.c_double),
        ("rv_mode", C.c_int), ("c_rv", C.c_double), ("c_uw", C.c_double),
        ("dx_avg", C.c_double), ("hv_first", C.c_int),
        ("ntags", C.c_int), ("tag_ptr", _i32p), ("tag_kind", _i32p),
        ("tag_fn", _i32p), ("tag_fn_params", _dp), ("bidx", _i32p),
        ("bnormal", _dp), ("btable", _dp),
        ("approx_du", _dp), ("success_iter", C.c_int),
        ("lap", _dp), ("tom_gamma", C.c_double), ("src_order", C.c_int * 4),
        ("igr_alpha", C.c_double), ("igr_maxiter", C.c_int), ("igr_sigma_out", _dp),
    ]


class StepCfg(C.Structure):
    _fields_ = [
        ("approx_order", C.c_int), ("limiter_on", C.c_int),
        ("lim_threshold", C.c_double * 2), ("lim_variable", C.c_int * 2),
        ("lim_count", C.c_int), ("ptr", _i64p), ("idx", _i32p),
    ]


def _declare(L):
    L.or_splitmix64.restype = C.c_uint64
    L.or_uniform01.restype = C.c_double
    L.or_cloud_square.argtypes = [C.c_int, C.c_uint64, C.c_double, _dp, _dp, _i8p, _dp, _dp]
    L.or_knn_brute.argtypes = [_dp, _dp, C.c_int64, C.c_int, _i32p, _dp]
    L.or_knn_grid.argtypes = [_dp, _dp, C.c_int64, C.c_int, _i32p, _dp]
    L.or_dx_stats.argtypes = [_dp, C.c_int64, _dp, _dp]
    L.or_bk_factor.argtypes = [C.c_int, _dp, C.c_int, _ip]
    L.or_bk_solve.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int, _dp, C.c_int]
    L.or_weights.argtypes = [_dp, _dp, C.c_int64, _i64p, _i32p, C.c_int, C.c_int, C.c_int,
                             _dp, _dp]
    L.or_weights_par.argtypes = [_dp, _dp, C.c_int64, _i64p, _i32p, C.c_int, C.c_int, C.c_int,
                                 _dp, _dp, C.c_int]
    L.or_phs_deriv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.or_phs_deriv.restype = C.c_double
    L.or_csc_build.argtypes = [C.c_int64, _i64p, _i32p, _i64p, _i32p, _i64p]
    L.or_csc_mul.argtypes = [C.c_int64, _i64p, _i32p, _dp, _dp, _dp, C.c_double, C.c_int]
    L.or_csc_tmul.argtypes = [C.c_int64, _i64p, _i32p, _dp, _dp, _dp, C.c_double]
    L.or_flux.argtypes = [C.c_int, _dp, C.c_double, C.c_int, _dp, _dp]
    L.or_bc_value.argtypes = [C.c_int, _dp, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_double, _dp]
    L.or_rhs.argtypes = [C.POINTER(RhsCfg), C.c_double, _dp, _dp, _dp, _i8p, _dp]
    L.or_rhs_par.argtypes = [C.POINTER(RhsCfg), _i64p, _i32p, _dp, _dp, C.c_double, _dp, _dp, _dp,
                             C.c_int]
    L.or_rhs_scratch_len.argtypes = [C.POINTER(RhsCfg)]
    L.or_rhs_scratch_len.restype = C.c_int64
    L.or_igr_rhs.argtypes = [C.POINTER(RhsCfg), _dp, _dp, _dp]
    L.or_igr_apply.argtypes = [C.POINTER(RhsCfg), _dp, _dp, _dp]
    L.or_igr_cg.argtypes = [C.POINTER(RhsCfg), _dp, _dp, _dp, C.c_int]
    L.or_igr_cg.restype = C.c_int
    L.or_rv_norms.argtypes = [C.c_int64, C.c_int, _dp, _dp, _dp]
    L.or_upwind_visc.argtypes = [C.POINTER(RhsCfg), _dp, _dp]
    L.or_update_visc.argtypes = [C.c_int64, _dp, _dp, C.c_int, _dp, _i8p]
    L.or_pairwise_sum.argtypes = [_dp, C.c_int64, C.c_int64]
    L.or_pairwise_sum.restype = C.c_double
    L.or_time_deriv_weights.argtypes = [C.c_int, _dp, _dp]
    L.or_history_update.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    L.or_limiter_zhang_shu.argtypes = [C.c_int64, C.c_int, C.c_double, _i64p, _i32p, C.c_double,
                                       C.c_int, _dp, _dp]
    L.or_ssprk33.argtypes = [C.POINTER(RhsCfg), C.POINTER(StepCfg), C.c_double, C.c_double,
                             C.c_int, _dp, _dp, _dp, _dp, _dp, _ip]
    L.or_rhs_vjp.argtypes = [C.POINTER(RhsCfg), C.c_double, _dp, _dp, _dp, _dp]
    L.or_npoly.argtypes = [C.c_int]


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else C.cast(None, t)


def d(a):
    return _p(a, _dp)


def i32(a):
    return _p(a, _i32p)


def i64(a):
    return _p(a, _i64p)


# ------------------------------------------------------------------ clouds
def cloud_square(s, seed, jitter=0.3):
    n = s * s
    x, y = np.empty(n), np.empty(n)
    tag = np.empty(n, np.int8)
    nx, ny = np.empty(n), np.empty(n)
    lib().or_cloud_square(s, C.c_uint64(seed), jitter, d(x), d(y), _p(tag, _i8p), d(nx), d(ny))
    return x, y, tag, nx, ny


# ---------------------------------------------------------------------- kNN
def knn_brute(x, y, k):
    n = len(x)
    idx = np.empty((n, k), np.int32)
    d2 = np.empty((n, k))
    lib().or_knn_brute(d(x), d(y), n, k, i32(idx), d(d2))
    return idx, d2


def knn_grid(x, y, k):
    n = len(x)
    idx = np.empty((n, k), np.int32)
    d2 = np.empty((n, k))
    rc = lib().or_knn_grid(d(np.ascontiguousarray(x)), d(np.ascontiguousarray(y)), n, k,
                           i32(idx), d(d2))
    assert rc == 0
    return idx, d2


def dx_stats(d2_nn):
    d2_nn = np.ascontiguousarray(d2_nn, dtype=np.float64)
    a, b = C.c_double(), C.c_double()
    lib().or_dx_stats(d(d2_nn), len(d2_nn), C.byref(a), C.byref(b))
    return a.value, b.value


# ------------------------------------------------------------------ weights
def bk_factor(A):
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    m = A.shape[0]
    ipiv = np.zeros(m, np.intc)
    info = lib().or_bk_factor(m, d(A), m, _p(ipiv, _ip))
    return A, ipiv, info


def bk_solve_full(A, B):
    F, ipiv, info = bk_factor(A)
    m = F.shape[0]
    Bf = np.asfortranarray(np.array(B, dtype=np.float64, ndmin=2).reshape(m, -1)).copy(order="F")
    lib().or_bk_solve(m, d(F), m, _p(ipiv, _ip), Bf.shape[1], d(Bf), m)
    return Bf, info


def npoly(deg):
    return (deg + 1) * (deg + 2) // 2


def phs_deriv(p, k, direction, x, y):
    return lib().or_phs_deriv(p, k, direction, x, y)


def weights(x, y, ptr, idx, phs_p, poly_deg, deriv_k=1, nthreads=None):
    """Ragged stencils: stencil e = idx[ptr[e]:ptr[e+1]] (centre first).
    nthreads: None runs the serial loop; otherwise OpenMP over chunks of 64
    stencils (0 = OpenMP's default thread count), the same bits."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.int32).ravel()
    n = len(ptr) - 1
    wx = np.zeros(len(idx))
    wy = np.zeros(len(idx))
    xs, ys = np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)
    if nthreads is None:
        rc = lib().or_weights(d(xs), d(ys), n, i64(ptr), i32(idx), phs_p, poly_deg, deriv_k, d(wx), d(wy))
    else:
        rc = lib().or_weights_par(d(xs), d(ys), n, i64(ptr), i32(idx), phs_p, poly_deg, deriv_k, d(wx),
                                  d(wy), int(nthreads))
    return wx, wy, rc


def ell_ptr(N, k):
    return np.arange(N + 1, dtype=np.int64) * k


# --------------------------------------------------------------- operators
class CSC:
    """CSC pattern of the rows-e / columns-idx operator, with the permutation
    from ELL/ragged storage into CSC order."""

    def __init__(self, N, ptr, idx):
        self.N = N
        self.ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        self.idx = np.ascontiguousarray(idx, dtype=np.int32).ravel()
        nnz = len(self.idx)
        self.colptr = np.zeros(N + 1, np.int64)
        self.rowval = np.zeros(nnz, np.int32)
        self.perm = np.zeros(nnz, np.int64)
        lib().or_csc_build(N, i64(self.ptr), i32(self.idx), i64(self.colptr), i32(self.rowval),
                           i64(self.perm))

    def values(self, w_ell):
        return np.ascontiguousarray(np.asarray(w_ell, dtype=np.float64).ravel()[self.perm])

    def mul(self, nz, x, out=None, alpha=1.0, overwrite=True):
        if out is None:
            out = np.zeros(self.N)
        lib().or_csc_mul(self.N, i64(self.colptr), i32(self.rowval), d(nz),
                         d(np.ascontiguousarray(x, dtype=np.float64)), d(out), alpha,
                         1 if overwrite else 0)
        return out

    def tmul(self, nz, x, out=None, alpha=1.0):
        if out is None:
            out = np.zeros(self.N)
        lib().or_csc_tmul(self.N, i64(self.colptr), i32(self.rowval), d(nz),
                          d(np.ascontiguousarray(x, dtype=np.float64)), d(out), alpha)
        return out


class Problem:
    """Oracle rhs! configuration.  All arrays are kept alive here.

    boundary: list of (kind, node_idx, normals (nb,2) or None, fn, params(8) or
    None, table (nb,nvar) or None) in the reference's NamedTuple order."""

    def __init__(self, x, y, ptr, idx, wx, wy, eq, nvar, gamma=1.4, adv=(1.0, 1.0),
                 boundary=(), hv=None, hv_gamma=0.0, rv_mode=RV_NONE, c_rv=1.0, c_uw=1.0,
                 dx_avg=0.0, hv_first=True, lap=None, tom_gamma=0.0, order=None, igr_alpha=0.0,
                 igr_maxiter=20):
        """lap, tom_gamma: Tominec source; igr_alpha, igr_maxiter: IGR source
        (sigma of the last rhs! in self.igr_sigma); order: the sources in
        keyword order (SRC_* values; default Flyer and viscosity by hv_first,
        then Tominec, then IGR)."""
        self.N = N = len(x)
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.csc = CSC(N, ptr, idx)
        self.dx = self.csc.values(wx)
        self.dy = self.csc.values(wy)
        # the row form (the weights' own ELL/ragged order), for rhs_par
        self.rwx = np.ascontiguousarray(np.asarray(wx, dtype=np.float64).ravel())
        self.rwy = np.ascontiguousarray(np.asarray(wy, dtype=np.float64).ravel())
        self.hv = self.csc.values(hv) if hv is not None else None
        self.lap = self.csc.values(lap) if lap is not None else None
        self.nvar = nvar
        self.approx_du = np.zeros(nvar * N)
        tag_ptr, kinds, fns, params, bidx, bnorm, btab = [0], [], [], [], [], [], []
        for (kind, nodes, normals, fn, prm, table) in boundary:
            nodes = np.asarray(nodes, dtype=np.int32)
            bidx.append(nodes)
            tag_ptr.append(tag_ptr[-1] + len(nodes))
            kinds.append(kind)
            fns.append(fn if fn is not None else FN_CONST)
            pp = np.zeros(8)
            if prm is not None:
                pp[:len(prm)] = prm
            params.append(pp)
            bnorm.append(np.zeros((len(nodes), 2)) if normals is None
                         else np.asarray(normals, dtype=np.float64).reshape(-1, 2))
            btab.append(np.zeros((len(nodes), nvar)) if table is None
                        else np.asarray(table, dtype=np.float64).reshape(-1, nvar))
        self.tag_ptr = np.asarray(tag_ptr, np.int32)
        self.tag_kind = np.asarray(kinds if kinds else [0], np.int32)
        self.tag_fn = np.asarray(fns if fns else [0], np.int32)
        self.tag_params = np.concatenate(params) if params else np.zeros(8)
        self.bidx = np.concatenate(bidx).astype(np.int32) if bidx else np.zeros(1, np.int32)
        self.bnormal = np.ascontiguousarray(np.concatenate(bnorm)) if bnorm else np.zeros(2)
        self.btable = np.ascontiguousarray(np.concatenate(btab)) if btab else np.zeros(nvar)
        c = RhsCfg()
        c.N, c.nvar, c.eq = N, nvar, eq
        c.adv_a[0], c.adv_a[1] = adv
        c.gamma = gamma
        c.px, c.py = d(self.x), d(self.y)
        c.colptr, c.rowval = i64(self.csc.colptr), i32(self.csc.rowval)
        c.dx, c.dy = d(self.dx), d(self.dy)
        c.hv = d(self.hv)
        c.hv_gamma = hv_gamma
        c.rv_mode, c.c_rv, c.c_uw, c.dx_avg = rv_mode, c_rv, c_uw, dx_avg
        c.hv_first = 1 if hv_first else 0
        c.ntags = len(kinds)
        c.tag_ptr, c.tag_kind, c.tag_fn = i32(self.tag_ptr), i32(self.tag_kind), i32(self.tag_fn)
        c.tag_fn_params = d(self.tag_params)
        c.bidx, c.bnormal, c.btable = i32(self.bidx), d(self.bnormal), d(self.btable)
        c.approx_du = d(self.approx_du)
        c.success_iter = 0
        c.lap = d(self.lap)
        c.tom_gamma = tom_gamma
        self.igr_sigma = np.zeros(N)
        c.igr_alpha, c.igr_maxiter, c.igr_sigma_out = igr_alpha, igr_maxiter, d(self.igr_sigma)
        for q, k in enumerate(order or []):
            c.src_order[q] = k
        self.cfg = c
        self.scratch = np.zeros(lib().or_rhs_scratch_len(C.byref(c)) + 3 * nvar * N)

    def igr_rhs(self, u):
        """(b, rinv) of the IGR system for state u (nvar, N)."""
        b, rinv = np.zeros(self.N), np.zeros(self.N)
        lib().or_igr_rhs(C.byref(self.cfg), d(np.ascontiguousarray(u, dtype=np.float64)), d(b), d(rinv))
        return b, rinv

    def igr_apply(self, rinv, v):
        y = np.zeros(self.N)
        lib().or_igr_apply(C.byref(self.cfg), d(np.ascontiguousarray(rinv)), d(np.ascontiguousarray(v)), d(y))
        return y

    def igr_cg(self, rinv, b, x0=None, maxiter=20):
        """(x, iterations) of cg! from x0 (zeros by default)."""
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        it = lib().or_igr_cg(C.byref(self.cfg), d(np.ascontiguousarray(rinv)), d(np.ascontiguousarray(b)), d(x),
                             maxiter)
        return x, it

    def rhs(self, u, t, success_iter=None, approx_du=None):
        """u: (nvar, N) array, mutated at boundary nodes like the reference.
        Returns du (nvar, N), eps, eps_c."""
        if approx_du is not None:
            self.approx_du[:] = np.asarray(approx_du).ravel()
        if success_iter is not None:
            self.cfg.success_iter = success_iter
        assert u.flags.c_contiguous and u.dtype == np.float64
        du = np.zeros_like(u)
        eps = np.zeros(self.N)
        eps_c = np.zeros(self.N, np.int8)
        lib().or_rhs(C.byref(self.cfg), t, d(u), d(du), d(eps), _p(eps_c, _i8p), d(self.scratch))
        return du, eps, eps_c

    def rhs_par(self, u, t, nthreads=0):
        """The CPU baseline's multi-threaded rhs! (OpenMP, or_rhs_par): du
        for the Euler equations with the viscosity sources only; agrees with
        rhs to rounding.  Mutates u at boundary nodes like rhs."""
        assert u.flags.c_contiguous and u.dtype == np.float64
        du = np.zeros_like(u)
        if not hasattr(self, "_par_scratch"):
            self._par_scratch = np.zeros(17 * self.N)
        rc = lib().or_rhs_par(C.byref(self.cfg), i64(self.csc.ptr), i32(self.csc.idx), d(self.rwx),
                              d(self.rwy), t, d(u), d(du), d(self._par_scratch), int(nthreads))
        if rc != 0:
            raise ValueError("rhs_par covers the Euler equations with viscosity sources only")
        return du

    def vjp(self, u, lam, t):
        ub = np.zeros_like(u)
        lib().or_rhs_vjp(C.byref(self.cfg), t, d(np.ascontiguousarray(u)),
                         d(np.ascontiguousarray(lam)), d(ub), d(self.scratch))
        return ub

    def ssprk33(self, u, t0, dt, nsteps, approx_order=0, state=None, limiter=None,
                stencil=None):
        """Runs nsteps of SSPRK33 in place on u (nvar, N).  state carries
        (fsal, time_hist, sol_hist, approx_du, success_iter) across calls."""
        nv = self.nvar * self.N
        P = max(approx_order, 0) + 1
        if state is None:
            state = dict(fsal=np.zeros(nv), th=np.zeros(P), sh=np.zeros(P * nv),
                         si=C.c_int(-1))
        s = StepCfg()
        s.approx_order = approx_order
        if limiter:
            s.limiter_on = 1
            s.lim_count = len(limiter)
            for q, (thr, var) in enumerate(limiter):
                s.lim_threshold[q] = thr
                s.lim_variable[q] = var
            ptr, idx = stencil
            state["lptr"] = np.ascontiguousarray(ptr, np.int64)
            state["lidx"] = np.ascontiguousarray(idx, np.int32).ravel()
            s.ptr, s.idx = i64(state["lptr"]), i32(state["lidx"])
        lib().or_ssprk33(C.byref(self.cfg), C.byref(s), t0, dt, nsteps, d(u), d(state["fsal"]),
                         d(state["th"]), d(state["sh"]), d(self.approx_du),
                         C.byref(state["si"]))
        return state


def time_deriv_weights(t):
    t = np.ascontiguousarray(t, dtype=np.float64)
    w = np.zeros(len(t))
    lib().or_time_deriv_weights(len(t), d(t), d(w))
    return w


def history_update(approx_order, success_iter, t, u, time_hist, sol_hist, approx_du):
    lib().or_history_update(len(u), approx_order, success_iter, t, d(u), d(time_hist),
                            d(sol_hist), d(approx_du))


def rv_norms(u):
    nvar, N = u.shape
    mean, ninf = np.zeros(nvar), np.zeros(nvar)
    lib().or_rv_norms(N, nvar, d(np.ascontiguousarray(u)), d(mean), d(ninf))
    return mean, ninf


def pairwise_sum(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().or_pairwise_sum(d(a), 0, len(a) - 1)


def limiter_zhang_shu(u, ptr, idx, threshold, variable, gamma=1.4):
    nvar, N = u.shape
    scratch = np.zeros(2 * nvar * N)
    ptr = np.ascontiguousarray(ptr, np.int64)
    idx = np.ascontiguousarray(idx, np.int32).ravel()
    lib().or_limiter_zhang_shu(N, nvar, gamma, i64(ptr), i32(idx), threshold, variable, d(u),
                               d(scratch))
    return u


def flux(eq, u, orientation, gamma=1.4, adv=(1.0, 1.0)):
    a = np.asarray(adv, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.zeros(len(u))
    lib().or_flux(eq, d(a), gamma, orientation, d(u), d(f))
    return f


def bc_value(fn, params, nvar, x, y, t, gamma=1.4):
    q = np.zeros(8)
    q[:len(params)] = params
    out = np.zeros(nvar)
    lib().or_bc_value(fn, d(q), nvar, gamma, x, y, t, d(out))
    return out
