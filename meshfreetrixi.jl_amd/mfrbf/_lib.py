"""ctypes binding of libmfrbf.so (C ABI in include/mfrbf.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfrbf.so")

# constants of include/mfrbf.h
MF_OK = 0
MF_ERR_ARGUMENT, MF_ERR_HIP, MF_ERR_STATE, MF_ERR_SINGULAR, MF_ERR_UNSUPPORTED = 1, 2, 3, 4, 5
MF_EQ_ADVECTION_2D, MF_EQ_EULER_2D = 0, 1
MF_BC_DIRICHLET, MF_BC_SLIP_WALL, MF_BC_DO_NOTHING = 0, 1, 2
MF_FN_CONST, MF_FN_PULSE, MF_FN_WAVE, MF_FN_GRADIENT, MF_FN_TABLE = 0, 1, 2, 3, 4
MF_OP_FLUX, MF_OP_HYPERVISCOSITY, MF_OP_LAPLACIAN = 0, 1, 2
MF_SRC_FLYER, MF_SRC_TOMINEC, MF_SRC_VISCOSITY, MF_SRC_IGR = 1, 2, 3, 4
MF_VISC_NONE, MF_VISC_RESIDUAL, MF_VISC_UPWIND = 0, 1, 2
MF_PREC_F64, MF_PREC_SPLIT32, MF_PREC_F64_REF = 0, 1, 2
MF_BUF_U, MF_BUF_DU, MF_BUF_AUX0, MF_BUF_AUX1 = 0, 1, 2, 3
MF_VAR_DENSITY, MF_VAR_PRESSURE = 0, 1
MF_STEP_GRAPH, MF_STEP_LIMITER, MF_STEP_HISTORY, MF_STEP_NOSYNC = 1, 2, 4, 8
MF_TUNE_FLUX_BATCH, MF_TUNE_TRANSPOSE_BATCH, MF_TUNE_NODE_ORDER, MF_TUNE_TILES, MF_TUNE_WEIGHTS_SOLVER = 0, 1, 2, 3, 4


class MfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libmfrbf error %d: %s" % (code, msg))
        self.code = code


_dp = C.POINTER(C.c_double)
_dpp = C.POINTER(_dp)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_i8p = C.POINTER(C.c_int8)
_ctx = C.c_void_p

_SIGS = {
    "mf_ctx_create": [C.c_int, C.POINTER(C.c_void_p)],
    "mf_ctx_destroy": [_ctx],
    "mf_synchronize": [_ctx],
    "mf_set_points": [_ctx, _dp, _dp, C.c_int64],
    "mf_set_points_grouped": [_ctx, _dp, _dp, C.c_int64, C.c_int, C.POINTER(C.c_int64)],
    "mf_knn": [_ctx, C.c_int, _i32p, _dp, _dp, _dp],
    "mf_knn_noalloc": [_ctx, C.c_int, C.c_void_p, C.c_void_p],
    "mf_set_stencil": [_ctx, _i32p, C.c_int64, C.c_int],
    "mf_compute_operator": [_ctx, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp],
    "mf_set_operator": [_ctx, C.c_int, _dp, _dp],
    "mf_weights_ragged_set": [_ctx, _i64p, _i32p, C.c_int64],
    "mf_weights_ragged_run": [_ctx, C.c_int, C.c_int, C.c_int],
    "mf_weights_ragged_get": [_ctx, _dp, _dp],
    "mf_set_equations": [_ctx, C.c_int, C.c_double, C.c_double],
    "mf_set_boundary": [_ctx, C.c_int, _i32p, _i32p, _i32p, _dp, _i32p, _dp],
    "mf_set_boundary_values": [_ctx, _dp],
    "mf_set_source_hyperviscosity": [_ctx, C.c_double],
    "mf_set_source_hyperviscosity_tominec": [_ctx, C.c_double],
    "mf_set_source_order": [_ctx, C.POINTER(C.c_int32), C.c_int],
    "mf_set_source_igr": [_ctx, C.c_double, C.c_int],
    "mf_get_igr_sigma": [_ctx, _dp],
    "mf_set_source_viscosity": [_ctx, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int],
    "mf_state_upload": [_ctx, C.c_int, _dpp],
    "mf_state_download": [_ctx, C.c_int, _dpp],
    "mf_state_device_ptr": [_ctx, C.c_int, C.POINTER(_dp)],
    "mf_rhs": [_ctx, C.c_double, C.c_int],
    "mf_rhs_host": [_ctx, C.c_double, C.c_int, _dpp, _dpp],
    "mf_get_viscosity": [_ctx, _dp, _i8p],
    "mf_get_viscosity_norms": [_ctx, _dp, _dp],
    "mf_history_update": [_ctx, C.c_double, C.c_int],
    "mf_get_approx_du": [_ctx, _dpp],
    "mf_limiter_zhang_shu": [_ctx, C.c_double, C.c_int],
    "mf_set_stage_limiter": [_ctx, C.c_int, _dp, _i32p],
    "mf_ssprk33": [_ctx, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int],
    "mf_rhs_vjp": [_ctx, C.c_double],
    "mf_ssprk33_vjp": [_ctx, C.c_double, C.c_double, C.c_int, C.c_int, _dp],
    "mf_time_rhs": [_ctx, C.c_double, C.c_int, C.c_int, _dp, _dp],
    "mf_time_rhs_cold": [_ctx, C.c_double, C.c_int, C.c_int, _dp],
    "mf_set_tuning": [_ctx, C.c_int, C.c_int, C.c_int],
    "mf_node_order": [_ctx, _i32p],
    "mf_tile_info": [_ctx, _i64p],
    "mf_state_copy": [_ctx, C.c_int, C.c_int],
    "mf_bandwidth_probe": [_ctx, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)],
}

_lib = None


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmfrbf.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.mf_last_error.argtypes = [_ctx]
        L.mf_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def exported_symbols():
    return sorted(list(_SIGS) + ["mf_last_error"])


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else C.cast(None, _dp)


def i32ptr(a):
    return a.ctypes.data_as(_i32p) if a is not None else C.cast(None, _i32p)


def i64ptr(a):
    return a.ctypes.data_as(_i64p) if a is not None else C.cast(None, _i64p)


def fields_ptr(u):
    """(nvar, N) C-contiguous float64 array -> double** of its rows."""
    arr = (_dp * u.shape[0])(*[u[v].ctypes.data_as(_dp) for v in range(u.shape[0])])
    return arr


class Context:
    """One libmfrbf context: a device, a stream and the device-resident data."""

    def __init__(self, device=0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.mf_ctx_create(device, C.byref(h))
        if rc != MF_OK:
            raise MfError(rc, "mf_ctx_create(device=%d) failed" % device)
        self.h = h
        self.device = device
        self.nvar = 1
        self.N = 0

    def close(self):
        if getattr(self, "h", None):
            self._L.mf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self._chk(self._L.mf_synchronize(self.h))

    def _chk(self, rc):
        if rc != MF_OK:
            msg = self._L.mf_last_error(self.h)
            raise MfError(rc, msg.decode() if msg else "")

    # ------------------------------------------------------------ cloud
    def set_points(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        self._chk(self._L.mf_set_points(self.h, dptr(x), dptr(y), len(x)))
        self.N = len(x)

    def set_points_grouped(self, x, y, group_ptr):
        """mf_set_points_grouped: the point set is len(group_ptr) - 1
        independent clouds, cloud g = points group_ptr[g] .. group_ptr[g+1]."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        gp = np.ascontiguousarray(group_ptr, dtype=np.int64)
        self._chk(self._L.mf_set_points_grouped(self.h, dptr(x), dptr(y), len(x), len(gp) - 1,
                                                gp.ctypes.data_as(C.POINTER(C.c_int64))))
        self.N = len(x)

    def knn(self, k, want_idx=True, want_dist=False):
        idx = np.empty((self.N, k), np.int32) if want_idx else None
        dist = np.empty((self.N, k)) if want_dist else None
        a, b = C.c_double(), C.c_double()
        self._chk(self._L.mf_knn(self.h, k, i32ptr(idx), dptr(dist), C.byref(a), C.byref(b)))
        return idx, dist, a.value, b.value

    def set_stencil(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        self._chk(self._L.mf_set_stencil(self.h, i32ptr(idx), idx.shape[0], idx.shape[1]))

    def knn_noalloc(self, k, d_idx, d_dist=None):
        """mf_knn_noalloc: stencils into caller-owned device memory (d_idx,
        d_dist are device addresses as integers; d_dist may be None)."""
        self._chk(self._L.mf_knn_noalloc(self.h, k, C.c_void_p(d_idx), C.c_void_p(d_dist) if d_dist else None))

    def compute_operator(self, p, deg, deriv, target, want=True, k=None):
        wx = wy = None
        if want:
            wx = np.empty((self.N, k))
            wy = np.empty((self.N, k))
        self._chk(self._L.mf_compute_operator(self.h, p, deg, deriv, target, dptr(wx), dptr(wy)))
        return wx, wy

    def set_operator(self, target, wx, wy=None):
        wx = np.ascontiguousarray(wx, dtype=np.float64)
        wy = None if wy is None else np.ascontiguousarray(wy, dtype=np.float64)
        self._chk(self._L.mf_set_operator(self.h, target, dptr(wx), dptr(wy)))

    def weights_ragged_set(self, ptr, idx):
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        self._chk(self._L.mf_weights_ragged_set(self.h, i64ptr(ptr), i32ptr(idx), len(ptr) - 1))
        self._rnnz = int(ptr[-1])

    def weights_ragged_run(self, p, deg, deriv=1):
        self._chk(self._L.mf_weights_ragged_run(self.h, p, deg, deriv))

    def weights_ragged_get(self):
        wx = np.empty(self._rnnz)
        wy = np.empty(self._rnnz)
        self._chk(self._L.mf_weights_ragged_get(self.h, dptr(wx), dptr(wy)))
        return wx, wy

    # ---------------------------------------------------------- physics
    def set_equations(self, eq, p0, p1=0.0):
        self._chk(self._L.mf_set_equations(self.h, eq, p0, p1))
        self.nvar = 4 if eq == MF_EQ_EULER_2D else 1

    def set_boundary(self, tag_ptr, kinds, fns, params, nodes, normals):
        tag_ptr = np.ascontiguousarray(tag_ptr, np.int32)
        kinds = np.ascontiguousarray(kinds, np.int32)
        fns = np.ascontiguousarray(fns, np.int32)
        params = np.ascontiguousarray(params, np.float64)
        nodes = np.ascontiguousarray(nodes, np.int32)
        normals = np.ascontiguousarray(normals, np.float64)
        self._chk(self._L.mf_set_boundary(self.h, len(kinds), i32ptr(tag_ptr), i32ptr(kinds), i32ptr(fns),
                                          dptr(params), i32ptr(nodes), dptr(normals)))

    def set_boundary_values(self, values):
        values = np.ascontiguousarray(values, np.float64)
        self._chk(self._L.mf_set_boundary_values(self.h, dptr(values)))

    def set_source_hyperviscosity(self, gamma):
        self._chk(self._L.mf_set_source_hyperviscosity(self.h, gamma))

    def set_source_hyperviscosity_tominec(self, gamma):
        self._chk(self._L.mf_set_source_hyperviscosity_tominec(self.h, gamma))

    def set_source_igr(self, alpha, maxiter=20):
        self._chk(self._L.mf_set_source_igr(self.h, alpha, maxiter))

    def igr_sigma(self):
        out = np.empty(self.N)
        self._chk(self._L.mf_get_igr_sigma(self.h, dptr(out)))
        return out

    def set_source_order(self, kinds):
        k = np.ascontiguousarray(kinds, np.int32)
        self._chk(self._L.mf_set_source_order(self.h, k.ctypes.data_as(C.POINTER(C.c_int32)), len(k)))

    def set_source_viscosity(self, mode, c_rv, c_uw, h, approx_order, hv_first):
        self._chk(self._L.mf_set_source_viscosity(self.h, mode, c_rv, c_uw, h, approx_order,
                                                  1 if hv_first else 0))

    # ------------------------------------------------------------ state
    def _fields(self, u):
        assert u.dtype == np.float64 and u.flags.c_contiguous and u.shape == (self.nvar, self.N), \
            "state must be a C-contiguous float64 (nvar, N) array"
        return fields_ptr(u)

    def upload(self, buf, u):
        self._chk(self._L.mf_state_upload(self.h, buf, self._fields(u)))

    def download(self, buf, out=None):
        if out is None:
            out = np.empty((self.nvar, self.N))
        self._chk(self._L.mf_state_download(self.h, buf, self._fields(out)))
        return out

    def rhs(self, t, precision=MF_PREC_F64):
        self._chk(self._L.mf_rhs(self.h, t, precision))

    def rhs_host(self, t, u, du, precision=MF_PREC_F64):
        self._chk(self._L.mf_rhs_host(self.h, t, precision, self._fields(u), self._fields(du)))

    def viscosity(self):
        eps = np.empty(self.N)
        code = np.empty(self.N, np.int8)
        self._chk(self._L.mf_get_viscosity(self.h, dptr(eps), code.ctypes.data_as(_i8p)))
        return eps, code

    def viscosity_norms(self):
        """(mean, ninf) of the last rhs! with the residual viscosity."""
        mean, ninf = np.empty(4), np.empty(4)
        self._chk(self._L.mf_get_viscosity_norms(self.h, dptr(mean), dptr(ninf)))
        return mean, ninf

    def history_update(self, t, success_iter):
        self._chk(self._L.mf_history_update(self.h, t, success_iter))

    def approx_du(self):
        out = np.empty((self.nvar, self.N))
        self._chk(self._L.mf_get_approx_du(self.h, self._fields(out)))
        return out

    def limiter_zhang_shu(self, threshold, variable):
        self._chk(self._L.mf_limiter_zhang_shu(self.h, threshold, variable))

    def set_stage_limiter(self, thresholds, variables):
        th = np.ascontiguousarray(thresholds, np.float64)
        va = np.ascontiguousarray(variables, np.int32)
        self._chk(self._L.mf_set_stage_limiter(self.h, len(th), dptr(th), i32ptr(va)))

    def ssprk33(self, t0, dt, nsteps, precision=MF_PREC_F64, flags=0, resume=False):
        self._chk(self._L.mf_ssprk33(self.h, t0, dt, nsteps, precision, flags, 1 if resume else 0))

    def rhs_vjp(self, t):
        self._chk(self._L.mf_rhs_vjp(self.h, t))

    def ssprk33_vjp(self, t0, dt, nsteps, checkpoint_every):
        loss = C.c_double()
        self._chk(self._L.mf_ssprk33_vjp(self.h, t0, dt, nsteps, checkpoint_every, C.byref(loss)))
        return loss.value

    def set_tuning(self, key, value, precision=-1):
        self._chk(self._L.mf_set_tuning(self.h, key, precision, value))

    def copy(self, dst_buf, src_buf):
        """Device-side copy of buffer src_buf into dst_buf."""
        self._chk(self._L.mf_state_copy(self.h, dst_buf, src_buf))

    def bandwidth_probe(self, kind, dst_buf, src_buf):
        """Time (ms) of a known-bytes kernel from src_buf into dst_buf: kind 0
        the streaming copy, 1 the windowed gather; kind 2 a read-modify-write
        over a 512 MiB scratch array (mf_bandwidth_probe)."""
        ms = C.c_double(0.0)
        self._chk(self._L.mf_bandwidth_probe(self.h, kind, dst_buf, src_buf, C.byref(ms)))
        return ms.value

    def tile_info(self):
        """dict of the LDS tile state (mf_tile_info)."""
        v = np.zeros(8, np.int64)
        self._chk(self._L.mf_tile_info(self.h, i64ptr(v)))
        keys = ("tiles_a", "tiles_b", "n_tiles", "halo_a_total", "halo_b_total", "halo_a_max", "halo_b_max",
                "b_entries_padded")
        return {k: int(x) for k, x in zip(keys, v)}

    def node_order(self):
        """order[p] = caller index of internal node p."""
        out = np.zeros(self.N, np.int32)
        self._chk(self._L.mf_node_order(self.h, i32ptr(out)))
        return out

    def time_rhs_cold(self, t, precision, n_calls):
        """Median ms of one rhs! after a cache flush (mf_time_rhs_cold)."""
        a = C.c_double()
        self._chk(self._L.mf_time_rhs_cold(self.h, t, precision, n_calls, C.byref(a)))
        return a.value

    def time_rhs(self, t, precision, n_calls):
        """(ms per rhs!, [ms of stencil kernel, transposed gather, reductions, boundary])."""
        a = C.c_double()
        parts = np.zeros(4)
        self._chk(self._L.mf_time_rhs(self.h, t, precision, n_calls, C.byref(a), dptr(parts)))
        return a.value, parts
