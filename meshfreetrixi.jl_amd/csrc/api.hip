// C ABI of libmfrbf.so (declared in include/mfrbf.h): context and buffer
// management, uploads/downloads, and the orchestration of the kernels in
// setup.hip, rhs.hip and adjoint.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>

#include "mf_device.h"

namespace mf {

int fail(mf_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_fail(mf_ctx* c, hipError_t e, const char* what)
{
    if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
    return MF_ERR_HIP;
}

template <class T>
int dalloc(mf_ctx* c, T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
    return MF_OK;
}
template int dalloc<double>(mf_ctx*, double**, size_t);
template int dalloc<double2>(mf_ctx*, double2**, size_t);
template int dalloc<float2>(mf_ctx*, float2**, size_t);
template int dalloc<float>(mf_ctx*, float**, size_t);
template int dalloc<int32_t>(mf_ctx*, int32_t**, size_t);
template int dalloc<int64_t>(mf_ctx*, int64_t**, size_t);
template int dalloc<int16_t>(mf_ctx*, int16_t**, size_t);
template int dalloc<int8_t>(mf_ctx*, int8_t**, size_t);
template int dalloc<uint8_t>(mf_ctx*, uint8_t**, size_t);
template int dalloc<uint16_t>(mf_ctx*, uint16_t**, size_t);
template int dalloc<uint64_t>(mf_ctx*, uint64_t**, size_t);
template int dalloc<IgrScalars>(mf_ctx*, IgrScalars**, size_t);

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// k-th derivative of r^p along s as sum_t c_t s^a_t r^b_t, built by
// repeated d/ds (s^a r^b) = a s^(a-1) r^b + b s^(a+1) r^(b-2).
PhsTable phs_table(int p, int k)
{
    PhsTable t{};
    double cc[16], c2[16];
    int aa[16], bb[16], a2[16], b2[16];
    int nt = 1;
    cc[0] = 1.0;
    aa[0] = 0;
    bb[0] = p;
    for (int it = 0; it < k; ++it) {
        int n2 = 0;
        for (int q = 0; q < nt; ++q) {
            if (aa[q] > 0 && n2 < 16) {
                c2[n2] = cc[q] * aa[q];
                a2[n2] = aa[q] - 1;
                b2[n2] = bb[q];
                ++n2;
            }
            if (bb[q] != 0 && n2 < 16) {
                c2[n2] = cc[q] * bb[q];
                a2[n2] = aa[q] + 1;
                b2[n2] = bb[q] - 2;
                ++n2;
            }
        }
        nt = 0;
        for (int q = 0; q < n2; ++q) {
            int f = -1;
            for (int r = 0; r < nt; ++r)
                if (aa[r] == a2[q] && bb[r] == b2[q]) f = r;
            if (f >= 0) {
                cc[f] += c2[q];
            } else {
                cc[nt] = c2[q];
                aa[nt] = a2[q];
                bb[nt] = b2[q];
                ++nt;
            }
        }
    }
    t.nt = nt;
    for (int q = 0; q < nt; ++q) {
        t.c[q] = cc[q];
        t.a[q] = aa[q];
        t.b[q] = bb[q];
    }
    return t;
}

int bc_pass_homog(mf_ctx* c, double* u);
int bc_pass_state(mf_ctx* c, double* u, double dt_off);

static void free_ops(mf_ctx* c)
{
    dfree(c->d_wxy);
    dfree(c->d_hv);
    dfree(c->d_wxy32);
    dfree(c->d_hv32);
    dfree(c->d_sell_off);
    dfree(c->d_sell_w);
    dfree(c->d_trow);
    dfree(c->d_tw);
    dfree(c->d_tw32);
    c->d_wxy = nullptr;
    c->d_hv = nullptr;
    c->d_wxy32 = nullptr;
    c->d_hv32 = nullptr;
    c->d_sell_off = nullptr;
    c->d_sell_w = nullptr;
    c->d_trow = nullptr;
    c->d_tw = nullptr;
    c->d_tw32 = nullptr;
    dfree(c->d_wx_s);
    dfree(c->d_wy_s);
    dfree(c->d_hv_s);
    c->d_wx_s = c->d_wy_s = c->d_hv_s = nullptr;
    dfree(c->d_lap);
    dfree(c->d_lap32);
    dfree(c->d_lap_s);
    dfree(c->d_tsrc);
    dfree(c->d_tl);
    dfree(c->d_tl32);
    c->d_lap = c->d_lap_s = c->d_tl = nullptr;
    c->d_lap32 = c->d_tl32 = nullptr;
    c->d_tsrc = nullptr;
    c->have_flux = c->have_hv = c->have_lap = c->have_T = c->lap_T = false;
}

static void free_graph(mf_ctx* c)
{
    if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
    if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
    if (c->multi_exec) (void)hipGraphExecDestroy(c->multi_exec);
    if (c->multi_graph) (void)hipGraphDestroy(c->multi_graph);
    c->step_exec = c->multi_exec = nullptr;
    c->step_graph = c->multi_graph = nullptr;
    c->graph_prec = -1;
}

static void free_state(mf_ctx* c)
{
    for (auto& b : c->d_buf) {
        dfree(b);
        b = nullptr;
    }
    dfree(c->d_S);
    dfree(c->d_T);
    dfree(c->d_G);
    dfree(c->d_du32);
    c->d_du32 = nullptr;
    dfree(c->d_L);
    c->d_L = nullptr;
    dfree(c->d_igr);
    dfree(c->d_igr_part);
    dfree(c->d_igr_sc);
    c->d_igr = nullptr;
    c->d_igr_part = nullptr;
    c->d_igr_sc = nullptr;
    dfree(c->d_eps);
    dfree(c->d_epsc);
    dfree(c->d_approx);
    dfree(c->d_hist);
    dfree(c->d_partials);
    dfree(c->d_stage_psum);
    dfree(c->d_lim);
    dfree(c->d_adj);
    c->d_S = c->d_T = c->d_G = c->d_eps = c->d_approx = c->d_hist = c->d_partials = c->d_lim = nullptr;
    c->d_stage_psum = nullptr;
    c->d_adj = nullptr;
    c->adj_len = 0;
    c->d_epsc = nullptr;
    c->fsal_valid = false;
    free_graph(c);
}

static void free_bc(mf_ctx* c)
{
    dfree(c->bc.bidx);
    dfree(c->bc.btag);
    dfree(c->bc.bnormal);
    dfree(c->bc.btable);
    dfree(c->bc.flag);
    dfree(c->bc.entry);
    dfree(c->bc.ecode);
    dfree(c->bc.d_kind);
    dfree(c->bc.d_kind_slip);
    dfree(c->bc.d_fn);
    dfree(c->bc.d_params);
    c->bc.d_kind = c->bc.d_kind_slip = c->bc.d_fn = nullptr;
    c->bc.d_params = nullptr;
    c->bc.bidx = c->bc.btag = c->bc.entry = c->bc.ecode = nullptr;
    c->bc.bnormal = c->bc.btable = nullptr;
    c->bc.flag = nullptr;
    c->have_bc = false;
}

static void free_stencil(mf_ctx* c)
{
    dfree(c->d_idx);
    dfree(c->d_didx);
    dfree(c->d_idx_s);
    dfree(c->d_perm_s);
    c->d_idx = nullptr;
    c->d_didx = nullptr;
    c->d_idx_s = nullptr;
    c->d_perm_s = nullptr;
    c->k = 0;
    c->didx_ok = false;
    free_tiles(c);
    free_ops(c);
}

// device buffers that depend on N and nvar
int ensure_state(mf_ctx* c)
{
    if (c->N == 0) return fail(c, MF_ERR_STATE, "no point cloud (mf_set_points)");
    const size_t nv = (size_t)c->N * c->nvar;
    int rc;
    for (auto& b : c->d_buf)
        if (!b && (rc = dalloc(c, &b, nv))) return rc;
    if (!c->d_S && (rc = dalloc(c, &c->d_S, nv))) return rc;
    if (!c->d_T && (rc = dalloc(c, &c->d_T, nv))) return rc;
    if (c->visc_mode != MF_VISC_NONE) {
        if (!c->d_G && (rc = dalloc(c, &c->d_G, 2 * nv))) return rc;
        if (!c->d_du32 && (rc = dalloc(c, &c->d_du32, nv))) return rc;
        if (!c->d_eps && (rc = dalloc(c, &c->d_eps, c->N))) return rc;
        if (!c->d_epsc && (rc = dalloc(c, &c->d_epsc, c->N))) return rc;
        if (!c->d_approx) {
            if ((rc = dalloc(c, &c->d_approx, nv))) return rc;
            MF_HIP(c, hipMemsetAsync(c->d_approx, 0, sizeof(double) * nv, c->stream));
        }
        if (!c->d_hist && (rc = dalloc(c, &c->d_hist, nv * (size_t)(c->approx_order + 1)))) return rc;
        if (!c->d_partials) {
            // reduction passes: the sum pass and the max pass up to kBlock
            // blocks of 4 nodes per thread (the max pass loads one sum partial
            // per thread, the stencil kernels one max partial per thread)
            c->n_psum = (int)std::min<int64_t>(kBlock, (c->N + 1023) / 1024);
            c->n_pmax = (int)std::min<int64_t>(kBlock, (c->N + 1023) / 1024);
            if ((rc = dalloc(c, &c->d_partials, 4 * (size_t)(c->n_psum + c->n_pmax)))) return rc;
        }
        if (!c->d_stage_psum && (rc = dalloc(c, &c->d_stage_psum, 4 * (size_t)((c->N + kBlock - 1) / kBlock))))
            return rc;
    }
    // limiter scratch: nvar N blended values, then N flag bytes
    if (!c->d_lim && (rc = dalloc(c, &c->d_lim, nv + (size_t)(c->N + 7) / 8))) return rc;
    return MF_OK;
}

int ensure_transpose(mf_ctx* c)
{
    if ((c->visc_mode == MF_VISC_NONE && c->tom_gamma == 0.0) || c->have_T) return MF_OK;
    if (!c->have_flux) return fail(c, MF_ERR_STATE, "no flux operator");
    return launch_build_transpose(c);
}

static int check_ctx(mf_ctx* c)
{
    if (!c) return MF_ERR_ARGUMENT;
    c->err.clear();
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    return MF_OK;
}

static int upload_fields(mf_ctx* c, double* d_dst, const double* const* fields)
{
    if (!fields) return fail(c, MF_ERR_ARGUMENT, "null fields");
    for (int v = 0; v < c->nvar; ++v)
        if (!fields[v]) return fail(c, MF_ERR_ARGUMENT, "null field pointer");
    // stage in d_T (SoA), then interleave
    for (int v = 0; v < c->nvar; ++v)
        MF_HIP(c, hipMemcpyAsync(c->d_T + (size_t)v * c->N, fields[v], sizeof(double) * c->N,
                                 hipMemcpyHostToDevice, c->stream));
    return launch_soa_to_aos(c, c->d_T, d_dst);
}

static int download_fields(mf_ctx* c, const double* d_src, double* const* fields)
{
    if (!fields) return fail(c, MF_ERR_ARGUMENT, "null fields");
    for (int v = 0; v < c->nvar; ++v)
        if (!fields[v]) return fail(c, MF_ERR_ARGUMENT, "null field pointer");
    int rc = launch_aos_to_soa(c, d_src, c->d_T);
    if (rc) return rc;
    for (int v = 0; v < c->nvar; ++v)
        MF_HIP(c, hipMemcpyAsync(fields[v], c->d_T + (size_t)v * c->N, sizeof(double) * c->N,
                                 hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

}  // namespace mf

using namespace mf;

extern "C" {

int mf_ctx_create(int device, mf_ctx** out)
{
    if (!out) return MF_ERR_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MF_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return MF_ERR_HIP;
    mf_ctx* c = new mf_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MF_ERR_HIP;
    }
    (void)hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
    if (dalloc(c, &c->d_sc, 1) != MF_OK ||
        hipMemset(c->d_sc, 0, sizeof(StepScalars)) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return MF_ERR_HIP;
    }
    *out = c;
    return MF_OK;
}

int mf_ctx_destroy(mf_ctx* c)
{
    if (!c) return MF_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_state(c);
    free_stencil(c);
    free_bc(c);
    dfree(c->d_x);
    dfree(c->d_y);
    dfree(c->d_uid);
    dfree(c->d_inv);
    dfree(c->d_sc);
    dfree(c->d_rptr);
    dfree(c->d_ridx);
    dfree(c->d_rorder);
    dfree(c->d_rwx);
    dfree(c->d_rwy);
    dfree(c->d_status);
    dfree(c->d_flush);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MF_OK;
}

const char* mf_last_error(const mf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mf_synchronize(mf_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

static int set_points(mf_ctx* c, const double* x, const double* y, int64_t n, int n_groups,
                      const int64_t* group_ptr);

int mf_set_points(mf_ctx* c, const double* x, const double* y, int64_t n)
{
    return set_points(c, x, y, n, 1, nullptr);
}

int mf_set_points_grouped(mf_ctx* c, const double* x, const double* y, int64_t n, int n_groups,
                          const int64_t* group_ptr)
{
    if (c && (n_groups < 1 || !group_ptr))
        return fail(c, MF_ERR_ARGUMENT, "mf_set_points_grouped: need n_groups >= 1 and group_ptr");
    return set_points(c, x, y, n, n_groups, group_ptr);
}

static int set_points(mf_ctx* c, const double* x, const double* y, int64_t n, int n_groups,
                      const int64_t* group_ptr)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!x || !y || n < 2) return fail(c, MF_ERR_ARGUMENT, "mf_set_points: need >= 2 points");
    if (n > (int64_t)INT32_MAX) return fail(c, MF_ERR_ARGUMENT, "mf_set_points: too many points");
    // the cell grid and the node order are built from the bounding box
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]))
            return fail(c, MF_ERR_ARGUMENT, "mf_set_points: non-finite coordinate");
    std::vector<int32_t> group;
    if (group_ptr) {
        if (group_ptr[0] != 0 || group_ptr[n_groups] != n)
            return fail(c, MF_ERR_ARGUMENT, "mf_set_points_grouped: group_ptr must run from 0 to n_points");
        group.resize(n);
        for (int g = 0; g < n_groups; ++g) {
            if (group_ptr[g + 1] < group_ptr[g])
                return fail(c, MF_ERR_ARGUMENT, "mf_set_points_grouped: group_ptr not monotone");
            for (int64_t i = group_ptr[g]; i < group_ptr[g + 1]; ++i) group[i] = g;
        }
    }
    MF_HIP(c, hipStreamSynchronize(c->stream));
    // a new cloud (or a new order) invalidates everything built on the old one
    free_state(c);
    free_stencil(c);
    free_bc(c);
    if (n != c->N) {
        dfree(c->d_x);
        dfree(c->d_y);
        dfree(c->d_uid);
        dfree(c->d_inv);
        c->d_x = c->d_y = nullptr;
        c->d_uid = c->d_inv = nullptr;
        c->N = n;
        if ((rc = dalloc(c, &c->d_x, n)) || (rc = dalloc(c, &c->d_y, n)) || (rc = dalloc(c, &c->d_uid, n)) ||
            (rc = dalloc(c, &c->d_inv, n)))
            return rc;
    }
    double *xu = nullptr, *yu = nullptr;
    int32_t* gu = nullptr;
    if ((rc = dalloc(c, &xu, n)) || (rc = dalloc(c, &yu, n))) return rc;
    if (group_ptr && (rc = dalloc(c, &gu, n))) return rc;
    MF_HIP(c, hipMemcpyAsync(xu, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemcpyAsync(yu, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (gu) MF_HIP(c, hipMemcpyAsync(gu, group.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    rc = launch_node_order(c, xu, yu, gu);
    (void)hipStreamSynchronize(c->stream);
    dfree(xu);
    dfree(yu);
    dfree(gu);
    c->n_groups = group_ptr ? n_groups : 1;
    if (group_ptr)
        c->group_ptr.assign(group_ptr, group_ptr + n_groups + 1);
    else
        c->group_ptr = {0, n};
    return rc;
}

static int set_k(mf_ctx* c, int k)
{
    free_stencil(c);
    free_graph(c);
    int rc;
    c->k = k;
    if ((rc = dalloc(c, &c->d_idx, (size_t)k * c->N)) || (rc = dalloc(c, &c->d_didx, (size_t)k * c->N)) ||
        (rc = dalloc(c, &c->d_idx_s, (size_t)k * c->N)) || (rc = dalloc(c, &c->d_perm_s, (size_t)k * c->N)))
        return rc;
    return MF_OK;
}

// derived stencil layouts once d_idx is filled
static int stencil_ready(mf_ctx* c)
{
    int rc;
    if ((rc = launch_build_didx(c))) return rc;
    return launch_sort_stencil(c);
}

int mf_knn(mf_ctx* c, int k, int32_t* idx_out, double* dist_out, double* dx_min, double* dx_avg)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->N == 0) return fail(c, MF_ERR_STATE, "mf_knn: no points");
    if (k < 2 || k > 64 || k > c->N) return fail(c, MF_ERR_ARGUMENT, "mf_knn: need 2 <= k <= min(64, N)");
    if (c->n_groups > 1)
        return fail(c, MF_ERR_UNSUPPORTED,
                    "mf_knn: the search would cross the clouds of a grouped set; search each cloud on its own");
    if ((rc = set_k(c, k))) return rc;
    double* d2 = nullptr;
    if ((rc = dalloc(c, &d2, (size_t)k * c->N))) return rc;
    if ((rc = launch_knn(c, k, d2))) return rc;
    if ((rc = stencil_ready(c))) return rc;
    double mn = 0, av = 0;
    if ((rc = launch_dx_stats(c, d2, &mn, &av))) return rc;
    if (dx_min) *dx_min = mn;
    if (dx_avg) *dx_avg = av;
    if (idx_out) {
        int32_t* rows = nullptr;
        if ((rc = dalloc(c, &rows, (size_t)k * c->N))) return rc;
        if ((rc = launch_ell_to_rows(c, c->d_idx, rows, k))) return rc;
        MF_HIP(c, hipMemcpyAsync(idx_out, rows, sizeof(int32_t) * k * c->N, hipMemcpyDeviceToHost, c->stream));
        MF_HIP(c, hipStreamSynchronize(c->stream));
        dfree(rows);
    }
    if (dist_out) {
        double* rows = nullptr;
        if ((rc = dalloc(c, &rows, (size_t)k * c->N))) return rc;
        if ((rc = launch_ell_to_rows_d(c, d2, rows, k))) return rc;
        MF_HIP(c, hipMemcpyAsync(dist_out, rows, sizeof(double) * k * c->N, hipMemcpyDeviceToHost, c->stream));
        MF_HIP(c, hipStreamSynchronize(c->stream));
        dfree(rows);
        for (int64_t q = 0; q < (int64_t)k * c->N; ++q) dist_out[q] = std::sqrt(dist_out[q]);
    }
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(d2);
    return MF_OK;
}

int mf_knn_noalloc(mf_ctx* c, int k, int32_t* d_idx_out, double* d_dist_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->N == 0) return fail(c, MF_ERR_STATE, "mf_knn_noalloc: no points");
    if (k < 2 || k > 64 || k > c->N) return fail(c, MF_ERR_ARGUMENT, "mf_knn_noalloc: need 2 <= k <= min(64, N)");
    if (!d_idx_out) return fail(c, MF_ERR_ARGUMENT, "mf_knn_noalloc: null output");
    if (c->n_groups > 1)
        return fail(c, MF_ERR_UNSUPPORTED,
                    "mf_knn_noalloc: the search would cross the clouds of a grouped set; search each cloud on its own");
    if (c->N > knn_noalloc_max_nodes())
        return fail(c, MF_ERR_UNSUPPORTED, "mf_knn_noalloc: more than 262,144 points, the size of its static scratch");
    return launch_knn_noalloc(c, k, d_idx_out, d_dist_out);
}

int mf_set_stencil(mf_ctx* c, const int32_t* idx, int64_t n_points, int k)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!idx || n_points != c->N || k < 1 || k > 64) return fail(c, MF_ERR_ARGUMENT, "mf_set_stencil: bad size");
    for (int64_t q = 0; q < n_points * k; ++q)
        if (idx[q] < 0 || idx[q] >= n_points) return fail(c, MF_ERR_ARGUMENT, "mf_set_stencil: index out of range");
    // a grouped set's stencils stay inside their node's cloud
    if (c->n_groups > 1)
        for (int g = 0; g < c->n_groups; ++g)
            for (int64_t i = c->group_ptr[g]; i < c->group_ptr[g + 1]; ++i)
                for (int s = 0; s < k; ++s) {
                    const int32_t j = idx[i * k + s];
                    if (j < c->group_ptr[g] || j >= c->group_ptr[g + 1])
                        return fail(c, MF_ERR_ARGUMENT, "mf_set_stencil: a stencil crosses the clouds of a grouped set");
                }
    if ((rc = set_k(c, k))) return rc;
    int32_t* rows = nullptr;
    if ((rc = dalloc(c, &rows, (size_t)k * c->N))) return rc;
    MF_HIP(c, hipMemcpyAsync(rows, idx, sizeof(int32_t) * k * c->N, hipMemcpyHostToDevice, c->stream));
    if ((rc = launch_rows_to_ell(c, rows, c->d_idx, k))) return rc;
    if ((rc = stencil_ready(c))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(rows);
    return MF_OK;
}

static int alloc_op(mf_ctx* c, int target)
{
    const size_t n = (size_t)c->N * c->k;
    int rc;
    if (target == MF_OP_FLUX) {
        if (!c->d_wxy && (rc = dalloc(c, &c->d_wxy, n))) return rc;
        if (!c->d_wx_s && (rc = dalloc(c, &c->d_wx_s, n))) return rc;
        if (!c->d_wy_s && (rc = dalloc(c, &c->d_wy_s, n))) return rc;
        if (!c->d_wxy32 && (rc = dalloc(c, &c->d_wxy32, n))) return rc;
    } else if (target == MF_OP_HYPERVISCOSITY) {
        if (!c->d_hv && (rc = dalloc(c, &c->d_hv, n))) return rc;
        if (!c->d_hv_s && (rc = dalloc(c, &c->d_hv_s, n))) return rc;
        if (!c->d_hv32 && (rc = dalloc(c, &c->d_hv32, n))) return rc;
    } else {
        if (!c->d_lap && (rc = dalloc(c, &c->d_lap, n))) return rc;
        if (!c->d_lap_s && (rc = dalloc(c, &c->d_lap_s, n))) return rc;
        if (!c->d_lap32 && (rc = dalloc(c, &c->d_lap32, n))) return rc;
    }
    return MF_OK;
}

static void op_installed(mf_ctx* c, int target)
{
    free_graph(c);
    if (target == MF_OP_FLUX) {
        c->have_flux = true;
        c->have_T = false;
        free_tiles_b(c);
    } else if (target == MF_OP_HYPERVISCOSITY) {
        c->have_hv = true;
    } else {
        c->have_lap = true;
        c->lap_T = false;
        c->lap_tb = false;
    }
}

// The null-space solve needs K = s Pi R Pi + QQ' positive definite: r^p
// (p odd) is conditionally positive definite of order ceil(p/2), which the
// monomials up to degree ceil(p/2) - 1 must span; other bases go to
// Bunch-Kaufman, which does not need it (as the reference's solve).
static bool ns_basis_ok(int p, int deg)
{
    return (p & 1) && deg <= 4 && deg >= (p + 1) / 2 - 1;
}

// The null-space solve forms Q by the Cholesky QR of P (n x np), which
// squares P's condition number; P is well conditioned only with enough
// nodes per monomial.  Stencils of fewer than 2 np nodes go to
// Bunch-Kaufman: at 11 nodes with the 10 cubic monomials (cond P up to
// 2e5) the null-space weights are off by 3e-7, at 20 nodes by 6e-13.
static bool ns_size_ok(int n, int np)
{
    return n >= 2 * np;
}

int mf_compute_operator(mf_ctx* c, int phs_power, int poly_degree, int deriv_order, int target,
                        double* wx_out, double* wy_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->k == 0) return fail(c, MF_ERR_STATE, "mf_compute_operator: no stencil");
    if (target != MF_OP_FLUX && target != MF_OP_HYPERVISCOSITY && target != MF_OP_LAPLACIAN)
        return fail(c, MF_ERR_ARGUMENT, "mf_compute_operator: bad target");
    if (phs_power < 1 || poly_degree < 0 || deriv_order < 1 || deriv_order > 8)
        return fail(c, MF_ERR_ARGUMENT, "mf_compute_operator: bad basis");
    const int np = (poly_degree + 1) * (poly_degree + 2) / 2;
    if (c->k + np > kMaxStencilM) return fail(c, MF_ERR_UNSUPPORTED, "mf_compute_operator: stencil + monomials > 64");
    if (np > c->k) return fail(c, MF_ERR_ARGUMENT, "mf_compute_operator: fewer stencil nodes than monomials");
    const size_t n = (size_t)c->N * c->k;
    double *wx = nullptr, *wy = nullptr, *rows = nullptr;
    int st = 0;
    // the work buffers are freed on every return (after the stream has
    // drained on the success path)
    auto body = [&]() -> int {
        int r;
        if ((r = dalloc(c, &wx, n)) || (r = dalloc(c, &wy, n))) return r;
        if (!c->d_status && (r = dalloc(c, &c->d_status, 1))) return r;
        MF_HIP(c, hipMemsetAsync(c->d_status, 0, sizeof(int), c->stream));
        if (c->weights_solver == 2 && ns_basis_ok(phs_power, poly_degree) && ns_size_ok(c->k, np))
            r = launch_weights_ns(c, nullptr, c->d_idx, c->k, nullptr, c->N, c->k, c->k, phs_power, poly_degree,
                                  deriv_order, wx, wy, c->d_status);
        else
            r = launch_weights(c, nullptr, c->d_idx, c->k, c->N, c->k, phs_power, poly_degree, deriv_order, wx, wy,
                               c->d_status);
        if (r) return r;
        if ((r = alloc_op(c, target))) return r;
        if ((r = launch_pack_ops(c, target, wx, wy))) return r;
        if ((r = launch_pack_sorted(c, target, wx, wy))) return r;
        MF_HIP(c, hipMemcpyAsync(&st, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        for (int w = 0; w < 2; ++w) {
            double* out = w == 0 ? wx_out : wy_out;
            if (!out) continue;
            if (!rows && (r = dalloc(c, &rows, n))) return r;
            if ((r = launch_ell_to_rows_d(c, w == 0 ? wx : wy, rows, c->k))) return r;
            MF_HIP(c, hipMemcpyAsync(out, rows, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
            MF_HIP(c, hipStreamSynchronize(c->stream));
        }
        MF_HIP(c, hipStreamSynchronize(c->stream));
        return MF_OK;
    };
    rc = body();
    if (rc) (void)hipStreamSynchronize(c->stream);  // nothing in flight reads them
    dfree(wx);
    dfree(wy);
    dfree(rows);
    if (rc) return rc;
    op_installed(c, target);
    if (st) return fail(c, MF_ERR_SINGULAR, "mf_compute_operator: singular stencil system");
    return MF_OK;
}

int mf_set_operator(mf_ctx* c, int target, const double* wx, const double* wy)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->k == 0) return fail(c, MF_ERR_STATE, "mf_set_operator: no stencil");
    if (!wx || (target == MF_OP_FLUX && !wy) || (target != MF_OP_FLUX && wy) ||
        (target != MF_OP_FLUX && target != MF_OP_HYPERVISCOSITY && target != MF_OP_LAPLACIAN))
        return fail(c, MF_ERR_ARGUMENT, "mf_set_operator: bad arguments");
    const size_t n = (size_t)c->N * c->k;
    double *rx = nullptr, *ex = nullptr, *ey = nullptr;
    // the staging buffers are freed on every return (after the stream has
    // drained on the success path)
    auto body = [&]() -> int {
        int r;
        if ((r = dalloc(c, &rx, n)) || (r = dalloc(c, &ex, n))) return r;
        if (wy && (r = dalloc(c, &ey, n))) return r;
        MF_HIP(c, hipMemcpyAsync(rx, wx, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        if ((r = launch_rows_to_ell_d(c, rx, ex, c->k))) return r;
        if (wy) {
            MF_HIP(c, hipStreamSynchronize(c->stream));
            MF_HIP(c, hipMemcpyAsync(rx, wy, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
            if ((r = launch_rows_to_ell_d(c, rx, ey, c->k))) return r;
        }
        if ((r = alloc_op(c, target))) return r;
        if ((r = launch_pack_ops(c, target, ex, ey))) return r;
        if ((r = launch_pack_sorted(c, target, ex, ey))) return r;
        MF_HIP(c, hipStreamSynchronize(c->stream));
        return MF_OK;
    };
    rc = body();
    if (rc) (void)hipStreamSynchronize(c->stream);  // nothing in flight reads them
    dfree(rx);
    dfree(ex);
    dfree(ey);
    if (rc) return rc;
    op_installed(c, target);
    return MF_OK;
}

int mf_weights_ragged_set(mf_ctx* c, const int64_t* ptr, const int32_t* idx, int64_t ns)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!ptr || !idx || ns < 1) return fail(c, MF_ERR_ARGUMENT, "mf_weights_ragged_set: bad arguments");
    if (c->N == 0) return fail(c, MF_ERR_STATE, "mf_weights_ragged_set: no points");
    int maxn = 0, minn = 64;
    for (int64_t e = 0; e < ns; ++e) {
        const int64_t n = ptr[e + 1] - ptr[e];
        if (n < 1 || n > 64) return fail(c, MF_ERR_ARGUMENT, "mf_weights_ragged_set: stencil size out of range");
        maxn = std::max<int>(maxn, (int)n);
        minn = std::min<int>(minn, (int)n);
    }
    const int64_t nnz = ptr[ns] - ptr[0];
    if (ptr[0] != 0) return fail(c, MF_ERR_ARGUMENT, "mf_weights_ragged_set: ptr[0] must be 0");
    for (int64_t q = 0; q < nnz; ++q)
        if (idx[q] < 0 || idx[q] >= c->N) return fail(c, MF_ERR_ARGUMENT, "mf_weights_ragged_set: index out of range");
    dfree(c->d_rptr);
    dfree(c->d_ridx);
    dfree(c->d_rorder);
    dfree(c->d_rwx);
    dfree(c->d_rwy);
    c->d_rptr = nullptr;
    c->d_ridx = nullptr;
    c->d_rorder = nullptr;
    c->d_rwx = c->d_rwy = nullptr;
    if ((rc = dalloc(c, &c->d_rptr, ns + 1)) || (rc = dalloc(c, &c->d_ridx, nnz)) ||
        (rc = dalloc(c, &c->d_rorder, ns)) ||
        (rc = dalloc(c, &c->d_rwx, nnz)) || (rc = dalloc(c, &c->d_rwy, nnz)))
        return rc;
    if (!c->d_status && (rc = dalloc(c, &c->d_status, 1))) return rc;
    std::vector<int32_t> ridx(nnz);
    for (int64_t q = 0; q < nnz; ++q) ridx[q] = c->h_inv[idx[q]];
    // group the stencils by size: one launch per size, LDS sized to it
    std::vector<int32_t> order(ns);
    c->rclass_start.clear();
    c->rclass_n.clear();
    {
        std::vector<int64_t> cnt(kMaxStencilM + 1, 0);
        for (int64_t e = 0; e < ns; ++e) ++cnt[ptr[e + 1] - ptr[e]];
        std::vector<int64_t> at(kMaxStencilM + 2, 0);
        for (int n = 1; n <= kMaxStencilM; ++n) at[n + 1] = at[n] + cnt[n];
        for (int n = 1; n <= kMaxStencilM; ++n)
            if (cnt[n]) {
                c->rclass_start.push_back(at[n]);
                c->rclass_n.push_back(n);
            }
        c->rclass_start.push_back(ns);
        for (int64_t e = 0; e < ns; ++e) order[at[ptr[e + 1] - ptr[e]]++] = (int32_t)e;
    }
    MF_HIP(c, hipMemcpyAsync(c->d_rptr, ptr, sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemcpyAsync(c->d_ridx, ridx.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemcpyAsync(c->d_rorder, order.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    c->rn = ns;
    c->rnnz = nnz;
    c->rmaxn = maxn;
    c->rminn = minn;
    return MF_OK;
}

int mf_weights_ragged_run(mf_ctx* c, int phs_power, int poly_degree, int deriv_order)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->rn == 0) return fail(c, MF_ERR_STATE, "mf_weights_ragged_run: no stencils");
    if (phs_power < 1 || poly_degree < 0 || deriv_order < 1 || deriv_order > 8)
        return fail(c, MF_ERR_ARGUMENT, "mf_weights_ragged_run: bad basis");
    MF_HIP(c, hipMemsetAsync(c->d_status, 0, sizeof(int), c->stream));
    const int np = (poly_degree + 1) * (poly_degree + 2) / 2;
    if (c->weights_solver != 1 && ns_basis_ok(phs_power, poly_degree)) {
        // the stencils in size order (d_rorder): the size classes below the
        // null-space size (they come first) by Bunch-Kaufman, one launch
        // each, then one launch per null-space variant: up to 32, 33 to 48
        // and 49 to 64 nodes
        size_t q0 = 0;
        for (; q0 < c->rclass_n.size() && !ns_size_ok(c->rclass_n[q0], np); ++q0) {
            const int64_t b0 = c->rclass_start[q0], b1 = c->rclass_start[q0 + 1];
            if ((rc = launch_weights(c, c->d_rptr, c->d_ridx, 0, b1 - b0, c->rclass_n[q0], phs_power, poly_degree,
                                     deriv_order, c->d_rwx, c->d_rwy, c->d_status, c->d_rorder, b0)))
                return rc;
        }
        int64_t lo = c->rclass_start[q0];
        for (int bound : {32, 48, 64}) {
            int64_t hi = lo;
            int maxn = 0, classes = 0;
            for (size_t q = q0; q < c->rclass_n.size(); ++q)
                if (c->rclass_n[q] <= bound) {
                    hi = c->rclass_start[q + 1];
                    maxn = c->rclass_n[q];
                    classes += c->rclass_start[q] >= lo;
                }
            if (hi > lo && (rc = launch_weights_ns(c, c->d_rptr, c->d_ridx, 0, c->d_rorder + lo, hi - lo, maxn,
                                                   classes == 1 ? maxn : 0, phs_power, poly_degree, deriv_order,
                                                   c->d_rwx, c->d_rwy, c->d_status)))
                return rc;
            lo = std::max(lo, hi);
        }
        return MF_OK;
    }
    for (size_t q = 0; q < c->rclass_n.size(); ++q) {
        const int64_t b0 = c->rclass_start[q], b1 = c->rclass_start[q + 1];
        if ((rc = launch_weights(c, c->d_rptr, c->d_ridx, 0, b1 - b0, c->rclass_n[q], phs_power, poly_degree,
                                 deriv_order, c->d_rwx, c->d_rwy, c->d_status, c->d_rorder, b0)))
            return rc;
    }
    return MF_OK;
}

int mf_weights_ragged_get(mf_ctx* c, double* wx, double* wy)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->rn == 0) return fail(c, MF_ERR_STATE, "mf_weights_ragged_get: no stencils");
    int st = 0;
    MF_HIP(c, hipMemcpyAsync(&st, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (wx) MF_HIP(c, hipMemcpyAsync(wx, c->d_rwx, sizeof(double) * c->rnnz, hipMemcpyDeviceToHost, c->stream));
    if (wy) MF_HIP(c, hipMemcpyAsync(wy, c->d_rwy, sizeof(double) * c->rnnz, hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    if (st) return fail(c, MF_ERR_SINGULAR, "ragged weights: singular stencil system");
    return MF_OK;
}

int mf_set_equations(mf_ctx* c, int equations, double p0, double p1)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    int nv;
    if (equations == MF_EQ_ADVECTION_2D) {
        nv = 1;
    } else if (equations == MF_EQ_EULER_2D) {
        nv = 4;
        if (!(p0 > 1.0)) return fail(c, MF_ERR_ARGUMENT, "mf_set_equations: gamma must exceed 1");
    } else {
        return fail(c, MF_ERR_ARGUMENT, "mf_set_equations: unknown equations");
    }
    MF_HIP(c, hipStreamSynchronize(c->stream));
    if (nv != c->nvar) {
        free_state(c);
        free_tiles(c);  // halo sizes that fit LDS depend on nvar
    }
    free_graph(c);
    c->eq = equations;
    c->nvar = nv;
    c->p0 = p0;
    c->p1 = p1;
    return MF_OK;
}

int mf_set_boundary(mf_ctx* c, int n_tags, const int32_t* tag_ptr, const int32_t* tag_kind,
                    const int32_t* tag_fn, const double* tag_params, const int32_t* node_idx,
                    const double* normals)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->N == 0) return fail(c, MF_ERR_STATE, "mf_set_boundary: no points");
    if (n_tags < 0 || n_tags > kMaxTags) return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary: too many tags");
    MF_HIP(c, hipStreamSynchronize(c->stream));
    free_bc(c);
    free_graph(c);
    BcDev& b = c->bc;
    b.ntags = n_tags;
    b.nb = 0;
    b.any_slip = 0;
    b.multi = 0;
    if (n_tags == 0) return MF_OK;
    b.tag_kind.assign(n_tags, 0);
    b.tag_fn.assign(n_tags, 0);
    b.tag_ptr.assign(n_tags + 1, 0);
    b.tag_params.assign(8 * (size_t)n_tags, 0.0);
    if (!tag_ptr || !tag_kind || !node_idx) return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary: null arrays");
    if (tag_ptr[0] != 0) return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary: tag_ptr[0] must be 0");
    // the whole of tag_ptr before any entry is read: a decreasing step
    // would otherwise index past the arrays sized by its last element
    for (int g = 0; g < n_tags; ++g)
        if (tag_ptr[g + 1] < tag_ptr[g]) return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary: tag_ptr not monotone");
    const int nb = tag_ptr[n_tags];
    std::vector<int32_t> btag(std::max(nb, 1));
    std::vector<int32_t> bnode(std::max(nb, 1));
    std::vector<uint8_t> flag(c->N, 0);
    std::vector<int32_t> entry(c->N, -1);
    for (int g = 0; g < n_tags; ++g) {
        const int kind = tag_kind[g];
        if (kind < MF_BC_DIRICHLET || kind > MF_BC_DO_NOTHING)
            return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary: unknown boundary kind");
        if (kind == MF_BC_SLIP_WALL && c->nvar != 4)
            return fail(c, MF_ERR_UNSUPPORTED, "mf_set_boundary: slip walls need the Euler equations");
        // the value function matters for Dirichlet tags only; the others
        // keep 0, so that it fits its 4 bits of the packed entry code
        const int fn = tag_fn && kind == MF_BC_DIRICHLET ? tag_fn[g] : MF_FN_CONST;
        if (kind == MF_BC_DIRICHLET && (fn < MF_FN_CONST || fn > MF_FN_TABLE))
            return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary: unknown value function");
        if (kind == MF_BC_DIRICHLET && (fn == MF_FN_WAVE || fn == MF_FN_GRADIENT) && c->nvar != 4)
            return fail(c, MF_ERR_UNSUPPORTED, "mf_set_boundary: Euler value function on a scalar equation");
        b.tag_kind[g] = kind;
        b.tag_fn[g] = fn;
        for (int q = 0; q < 8; ++q) b.tag_params[8 * g + q] = tag_params ? tag_params[8 * g + q] : 0.0;
        b.tag_ptr[g] = tag_ptr[g];
        if (kind == MF_BC_SLIP_WALL) b.any_slip = 1;
        for (int e = tag_ptr[g]; e < tag_ptr[g + 1]; ++e) {
            if (node_idx[e] < 0 || node_idx[e] >= c->N)
                return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary: node index out of range");
            const int32_t node = c->h_inv[node_idx[e]];
            bnode[e] = node;
            btag[e] = g;
            if (entry[node] >= 0) b.multi = 1;
            entry[node] = e;
            if (kind == MF_BC_DIRICHLET) flag[node] |= 1u;
            if (kind == MF_BC_SLIP_WALL) flag[node] |= 2u;
        }
    }
    b.tag_ptr[n_tags] = nb;
    b.nb = nb;
    if ((rc = dalloc(c, &b.bidx, nb)) || (rc = dalloc(c, &b.btag, nb)) || (rc = dalloc(c, &b.bnormal, 2 * (size_t)nb)) ||
        (rc = dalloc(c, &b.btable, (size_t)nb * 4)) || (rc = dalloc(c, &b.flag, c->N)) || (rc = dalloc(c, &b.entry, c->N)) ||
        (rc = dalloc(c, &b.ecode, std::max(nb, 1))) ||
        (rc = dalloc(c, &b.d_kind, n_tags)) || (rc = dalloc(c, &b.d_kind_slip, n_tags)) ||
        (rc = dalloc(c, &b.d_fn, n_tags)) || (rc = dalloc(c, &b.d_params, 8 * (size_t)n_tags)))
        return rc;
    std::vector<int32_t> kind_slip(n_tags);
    for (int g = 0; g < n_tags; ++g)
        kind_slip[g] = b.tag_kind[g] == MF_BC_SLIP_WALL ? MF_BC_SLIP_WALL : MF_BC_DO_NOTHING;
    MF_HIP(c, hipMemcpyAsync(b.d_kind, b.tag_kind.data(), sizeof(int32_t) * n_tags, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemcpyAsync(b.d_kind_slip, kind_slip.data(), sizeof(int32_t) * n_tags, hipMemcpyHostToDevice,
                             c->stream));
    MF_HIP(c, hipMemcpyAsync(b.d_fn, b.tag_fn.data(), sizeof(int32_t) * n_tags, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemcpyAsync(b.d_params, b.tag_params.data(), sizeof(double) * 8 * n_tags, hipMemcpyHostToDevice,
                             c->stream));
    std::vector<double> nrm(2 * (size_t)std::max(nb, 1), 0.0);
    if (normals) std::copy(normals, normals + 2 * (size_t)nb, nrm.begin());
    MF_HIP(c, hipMemcpyAsync(b.bidx, bnode.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemcpyAsync(b.btag, btag.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, c->stream));
    // each entry's kind, value function and tag in one word: the fused
    // first pass (sum pass, stage hand-over) reads them with one load
    std::vector<int32_t> ecode(std::max(nb, 1), 0);
    for (int e = 0; e < nb; ++e)
        ecode[e] = b.tag_kind[btag[e]] | (b.tag_fn[btag[e]] << 4) | (btag[e] << 8);
    MF_HIP(c, hipMemcpyAsync(b.ecode, ecode.data(), sizeof(int32_t) * ecode.size(), hipMemcpyHostToDevice,
                             c->stream));
    MF_HIP(c, hipMemcpyAsync(b.bnormal, nrm.data(), sizeof(double) * 2 * nb, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemsetAsync(b.btable, 0, sizeof(double) * nb * 4, c->stream));
    MF_HIP(c, hipMemcpyAsync(b.flag, flag.data(), c->N, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipMemcpyAsync(b.entry, entry.data(), sizeof(int32_t) * c->N, hipMemcpyHostToDevice, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    c->have_bc = true;
    return MF_OK;
}

int mf_set_boundary_values(mf_ctx* c, const double* values)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->have_bc) return fail(c, MF_ERR_STATE, "mf_set_boundary_values: no boundary");
    if (!values) return fail(c, MF_ERR_ARGUMENT, "mf_set_boundary_values: null values");
    MF_HIP(c, hipMemcpyAsync(c->bc.btable, values, sizeof(double) * c->bc.nb * c->nvar, hipMemcpyHostToDevice,
                             c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

int mf_set_source_hyperviscosity(mf_ctx* c, double gamma)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (gamma != 0.0 && !c->have_hv) return fail(c, MF_ERR_STATE, "hyperviscosity source without its operator");
    free_graph(c);
    c->hv_gamma = gamma;
    return MF_OK;
}

int mf_set_source_viscosity(mf_ctx* c, int mode, double c_rv, double c_uw, double h, int approx_order,
                            int hv_first)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (mode < MF_VISC_NONE || mode > MF_VISC_UPWIND) return fail(c, MF_ERR_ARGUMENT, "unknown viscosity mode");
    if (mode != MF_VISC_NONE && c->eq != MF_EQ_EULER_2D)
        return fail(c, MF_ERR_UNSUPPORTED, "viscosity sources need the Euler equations");
    if (approx_order < 0 || approx_order + 1 > kMaxHist) return fail(c, MF_ERR_ARGUMENT, "approx_order out of range");
    MF_HIP(c, hipStreamSynchronize(c->stream));
    free_graph(c);
    dfree(c->d_hist);
    c->d_hist = nullptr;
    c->visc_mode = mode;
    c->c_rv = c_rv;
    c->c_uw = c_uw;
    c->h_visc = h;
    c->approx_order = approx_order;
    // hv_first orders the Flyer and viscosity sources only
    int* rk = c->src_rank;
    if ((rk[MF_SRC_FLYER] < rk[MF_SRC_VISCOSITY]) != (hv_first != 0)) std::swap(rk[MF_SRC_FLYER], rk[MF_SRC_VISCOSITY]);
    c->hv_first = hv_first ? 1 : 0;
    return MF_OK;
}

int mf_set_source_hyperviscosity_tominec(mf_ctx* c, double gamma)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (gamma != 0.0 && !c->have_lap)
        return fail(c, MF_ERR_STATE, "Tominec hyperviscosity without its operator (MF_OP_LAPLACIAN)");
    if (gamma != 0.0 && c->eq != MF_EQ_EULER_2D)
        return fail(c, MF_ERR_UNSUPPORTED, "Tominec hyperviscosity needs the Euler equations");
    free_graph(c);
    c->tom_gamma = gamma;
    return MF_OK;
}

int mf_set_source_igr(mf_ctx* c, double alpha, int maxiter)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (maxiter < 0 || maxiter > 10000) return fail(c, MF_ERR_ARGUMENT, "mf_set_source_igr: maxiter out of range");
    if (alpha != 0.0 && c->eq != MF_EQ_EULER_2D)
        return fail(c, MF_ERR_UNSUPPORTED, "the IGR source needs the Euler equations");
    free_graph(c);
    c->igr_alpha = alpha;
    c->igr_maxiter = maxiter;
    return MF_OK;
}

int mf_get_igr_sigma(mf_ctx* c, double* sigma)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!sigma) return fail(c, MF_ERR_ARGUMENT, "mf_get_igr_sigma: null output");
    if (!c->d_igr) return fail(c, MF_ERR_STATE, "mf_get_igr_sigma: no rhs! with the IGR source yet");
    double* tmp = nullptr;
    if ((rc = dalloc(c, &tmp, c->N))) return rc;
    if ((rc = launch_gather_nodes_d(c, c->d_igr + 2 * c->N, tmp))) return rc;
    MF_HIP(c, hipMemcpyAsync(sigma, tmp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(tmp);
    return MF_OK;
}

int mf_set_source_order(mf_ctx* c, const int32_t* kinds, int n)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 0 || n > 4 || (n > 0 && !kinds)) return fail(c, MF_ERR_ARGUMENT, "mf_set_source_order: bad list");
    bool seen[5] = {false, false, false, false, false};
    for (int q = 0; q < n; ++q) {
        if (kinds[q] < MF_SRC_FLYER || kinds[q] > MF_SRC_IGR || seen[kinds[q]])
            return fail(c, MF_ERR_ARGUMENT, "mf_set_source_order: kinds must be distinct MF_SRC_* values");
        seen[kinds[q]] = true;
    }
    free_graph(c);
    int r = 0;
    for (int q = 0; q < n; ++q) c->src_rank[kinds[q]] = r++;
    for (int k = MF_SRC_FLYER; k <= MF_SRC_IGR; ++k)
        if (!seen[k]) c->src_rank[k] = r++;
    c->hv_first = c->src_rank[MF_SRC_FLYER] < c->src_rank[MF_SRC_VISCOSITY] ? 1 : 0;
    return MF_OK;
}

int mf_state_upload(mf_ctx* c, int buf, const double* const* fields)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (buf < 0 || buf > 3) return fail(c, MF_ERR_ARGUMENT, "bad buffer");
    if ((rc = ensure_state(c))) return rc;
    if ((rc = upload_fields(c, c->d_buf[buf], fields))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    if (buf == MF_BUF_U) c->fsal_valid = false;
    return MF_OK;
}

int mf_state_download(mf_ctx* c, int buf, double* const* fields)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (buf < 0 || buf > 3) return fail(c, MF_ERR_ARGUMENT, "bad buffer");
    if ((rc = ensure_state(c))) return rc;
    return download_fields(c, c->d_buf[buf], fields);
}

int mf_state_device_ptr(mf_ctx* c, int buf, double** dptr)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (buf < 0 || buf > 3 || !dptr) return fail(c, MF_ERR_ARGUMENT, "bad buffer");
    if ((rc = ensure_state(c))) return rc;
    *dptr = c->d_buf[buf];
    return MF_OK;
}

int mf_state_copy(mf_ctx* c, int dst_buf, int src_buf)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (dst_buf < 0 || dst_buf > 3 || src_buf < 0 || src_buf > 3) return fail(c, MF_ERR_ARGUMENT, "bad buffer");
    if ((rc = ensure_state(c))) return rc;
    if (dst_buf == src_buf) return MF_OK;
    if ((rc = launch_copy_buffer(c, c->d_buf[src_buf], c->d_buf[dst_buf], c->N * c->nvar))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    if (dst_buf == MF_BUF_U) c->fsal_valid = false;
    return MF_OK;
}

int mf_bandwidth_probe(mf_ctx* c, int kind, int dst_buf, int src_buf, double* ms)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (dst_buf < 0 || dst_buf > 3 || src_buf < 0 || src_buf > 3 || dst_buf == src_buf || kind < 0 || kind > 2)
        return fail(c, MF_ERR_ARGUMENT, "mf_bandwidth_probe: bad arguments");
    if ((rc = ensure_state(c))) return rc;
    if (kind == 2 && !c->d_flush && (rc = dalloc(c, &c->d_flush, kFlushDoubles))) return rc;
    hipEvent_t e0, e1;
    MF_HIP(c, hipEventCreate(&e0));
    MF_HIP(c, hipEventCreate(&e1));
    MF_HIP(c, hipEventRecord(e0, c->stream));
    const int64_t n = c->N * c->nvar;
    rc = kind == 0   ? launch_copy_buffer(c, c->d_buf[src_buf], c->d_buf[dst_buf], n)
         : kind == 1 ? launch_gather_probe(c, c->d_buf[src_buf], c->d_buf[dst_buf], n)
                     : launch_flush(c, c->d_flush, kFlushDoubles);
    if (rc == MF_OK) {
        MF_HIP(c, hipEventRecord(e1, c->stream));
        MF_HIP(c, hipEventSynchronize(e1));
        float t = 0.0f;
        MF_HIP(c, hipEventElapsedTime(&t, e0, e1));
        if (ms) *ms = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (dst_buf == MF_BUF_U) c->fsal_valid = false;
    return rc;
}

int mf_node_order(mf_ctx* c, int32_t* order_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->N == 0) return fail(c, MF_ERR_STATE, "mf_node_order: no points");
    if (!order_out) return fail(c, MF_ERR_ARGUMENT, "mf_node_order: null output");
    MF_HIP(c, hipMemcpyAsync(order_out, c->d_uid, sizeof(int32_t) * c->N, hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

static int prepare_rhs(mf_ctx* c)
{
    int rc;
    if ((rc = ensure_state(c))) return rc;
    if ((rc = ensure_transpose(c))) return rc;
    if (c->hv_gamma != 0.0 && !c->have_hv) return fail(c, MF_ERR_STATE, "hyperviscosity operator missing");
    const bool rv = c->visc_mode != MF_VISC_NONE;
    const bool tom = c->tom_gamma != 0.0;
    if (tom && !c->have_lap) return fail(c, MF_ERR_STATE, "Tominec operator missing");
    if (tom) {
        if (!c->d_L && (rc = dalloc(c, &c->d_L, (size_t)c->N * c->nvar))) return rc;
        if (!c->d_du32 && (rc = dalloc(c, &c->d_du32, (size_t)c->N * c->nvar))) return rc;
        if (!c->lap_T) {
            free_graph(c);
            if ((!c->d_tl && (rc = dalloc(c, &c->d_tl, c->t_entries))) ||
                (!c->d_tl32 && (rc = dalloc(c, &c->d_tl32, c->t_entries))))
                return rc;
            if ((rc = launch_fill_lap_t(c, c->t_entries, c->d_tsrc, c->d_tl, c->d_tl32))) return rc;
            MF_HIP(c, hipStreamSynchronize(c->stream));
            c->lap_T = true;
        }
    }
    if (c->igr_alpha != 0.0) {
        const size_t nblk = (size_t)((c->N + 255) / 256);
        if ((!c->d_igr && (rc = dalloc(c, &c->d_igr, 9 * (size_t)c->N))) ||
            (!c->d_igr_part && (rc = dalloc(c, &c->d_igr_part, 2 * nblk))) ||
            (!c->d_igr_sc && (rc = dalloc(c, &c->d_igr_sc, 1))))
            return rc;
    }
    const bool want_b = (rv || tom) && c->nvar == 4;
    // phase-B halo record: G (2 nvar doubles) and / or lap u (nvar)
    const int rec_b = (rv ? 2 * c->nvar : 0) + (tom ? c->nvar : 0);
    if (want_b && c->tried_b && c->tiles_b_rec != rec_b) {
        free_graph(c);
        free_tiles_b(c);
    }
    if (c->use_tiles && c->have_flux && (!c->tried_a || (want_b && c->tiles_a && !c->tried_b))) {
        free_graph(c);
        if ((rc = build_tiles(c, want_b, rec_b))) return rc;
    }
    if (c->tiles_a && c->tm_dirty) {
        free_graph(c);
        if ((rc = ensure_tile_ops(c))) return rc;
        MF_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (tom && c->tiles_b && !c->lap_tb) {
        free_graph(c);
        if ((!c->d_twl && (rc = dalloc(c, &c->d_twl, c->nb_entries))) ||
            (!c->d_twl32 && (rc = dalloc(c, &c->d_twl32, c->nb_entries))))
            return rc;
        if ((rc = launch_fill_lap_t(c, c->nb_entries, c->d_tb_src, c->d_twl, c->d_twl32))) return rc;
        MF_HIP(c, hipStreamSynchronize(c->stream));
        c->lap_tb = true;
    }
    return MF_OK;
}

int mf_rhs(mf_ctx* c, double t, int precision)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = prepare_rhs(c))) return rc;
    if ((rc = run_set_time(c, t, -1))) return rc;
    StageOut none;
    if ((rc = run_rhs(c, precision, c->d_buf[MF_BUF_U], c->d_buf[MF_BUF_DU], 0.0, none))) return rc;
    c->fsal_valid = false;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

int mf_rhs_host(mf_ctx* c, double t, int precision, double* const* u, double* const* du)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = prepare_rhs(c))) return rc;
    if ((rc = upload_fields(c, c->d_buf[MF_BUF_U], u))) return rc;
    if ((rc = mf_rhs(c, t, precision))) return rc;
    if ((rc = download_fields(c, c->d_buf[MF_BUF_DU], du))) return rc;
    return download_fields(c, c->d_buf[MF_BUF_U], u);
}

int mf_get_viscosity(mf_ctx* c, double* eps, int8_t* eps_c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->visc_mode == MF_VISC_NONE || !c->d_eps) return fail(c, MF_ERR_STATE, "no viscosity source evaluated");
    double* tmp = nullptr;
    if ((rc = dalloc(c, &tmp, c->N))) return rc;
    if (eps) {
        if ((rc = launch_gather_nodes_d(c, c->d_eps, tmp))) return rc;
        MF_HIP(c, hipMemcpyAsync(eps, tmp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
        MF_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (eps_c) {
        int8_t* t8 = reinterpret_cast<int8_t*>(tmp);
        if ((rc = launch_gather_nodes_i8(c, c->d_epsc, t8))) return rc;
        MF_HIP(c, hipMemcpyAsync(eps_c, t8, c->N, hipMemcpyDeviceToHost, c->stream));
    }
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(tmp);
    return MF_OK;
}

int mf_get_viscosity_norms(mf_ctx* c, double* mean, double* ninf)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->visc_mode != MF_VISC_RESIDUAL || !c->d_partials)
        return fail(c, MF_ERR_STATE, "no residual viscosity evaluated");
    StepScalars sc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    MF_HIP(c, hipMemcpy(&sc, c->d_sc, sizeof(StepScalars), hipMemcpyDeviceToHost));
    for (int v = 0; v < 4; ++v) {
        if (mean) mean[v] = sc.norms[v];
        if (ninf) ninf[v] = sc.norms[4 + v];
    }
    return MF_OK;
}

int mf_history_update(mf_ctx* c, double t, int success_iter)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->visc_mode != MF_VISC_RESIDUAL) return MF_OK;  // modify_cache! fallback: no-op
    if (success_iter < 0) return fail(c, MF_ERR_ARGUMENT, "success_iter < 0");
    if ((rc = ensure_state(c))) return rc;
    if ((rc = run_set_time(c, t, success_iter))) return rc;
    if ((rc = run_history(c, c->d_buf[MF_BUF_U], success_iter == 0 ? 1 : 0))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    c->fsal_valid = false;
    return MF_OK;
}

int mf_get_approx_du(mf_ctx* c, double* const* fields)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->d_approx) return fail(c, MF_ERR_STATE, "no residual viscosity history");
    return download_fields(c, c->d_approx, fields);
}

int mf_limiter_zhang_shu(mf_ctx* c, double threshold, int variable)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->k == 0) return fail(c, MF_ERR_STATE, "limiter: no stencil");
    if (variable != MF_VAR_DENSITY && variable != MF_VAR_PRESSURE) return fail(c, MF_ERR_ARGUMENT, "limiter: bad variable");
    if (variable == MF_VAR_PRESSURE && c->nvar != 4) return fail(c, MF_ERR_UNSUPPORTED, "limiter: pressure needs Euler");
    if ((rc = ensure_state(c))) return rc;
    if ((rc = run_limiter(c, c->d_buf[MF_BUF_U], threshold, variable))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

int mf_set_stage_limiter(mf_ctx* c, int n, const double* thresholds, const int32_t* variables)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 0 || n > 2 || (n > 0 && (!thresholds || !variables))) return fail(c, MF_ERR_ARGUMENT, "stage limiter: bad arguments");
    for (int q = 0; q < n; ++q)
        if (variables[q] != MF_VAR_DENSITY && variables[q] != MF_VAR_PRESSURE)
            return fail(c, MF_ERR_ARGUMENT, "stage limiter: bad variable");
    free_graph(c);
    c->lim_n = n;
    for (int q = 0; q < n; ++q) {
        c->lim_thr[q] = thresholds[q];
        c->lim_var[q] = variables[q];
    }
    return MF_OK;
}

// Whether the stages of mf_ssprk33 hand each other the next rhs!'s first
// boundary pass and sums (the residual viscosity's transposed kernel
// writes each stage; no stage limiter runs between them).
static bool stage_sums(const mf_ctx* c, bool limiter)
{
    return c->visc_mode == MF_VISC_RESIDUAL && !limiter && c->igr_alpha == 0.0 && c->d_stage_psum &&
           !(c->have_bc && c->bc.multi) && !(c->tiles_b && c->n_tiles != (c->N + kBlock - 1) / kBlock);
}

// One SSPRK33 step on U with the FSAL stage S = U + dt f(U) already formed.
// With sums, every rhs! of the step takes its input's boundary pass and
// sums from the rhs! before it and leaves them for the next: T's and U's
// at this step's stage times t + dt/2 and t + dt, S's at the next step's
// t + dt (the time is advanced after the last rhs!).
static int ssprk_step(mf_ctx* c, double dt, int prec, bool limiter, bool history, bool sums)
{
    int rc;
    auto limit = [&](double* u) -> int {
        if (!limiter) return MF_OK;
        for (int q = 0; q < c->lim_n; ++q) {
            int r2 = run_limiter(c, u, c->lim_thr[q], c->lim_var[q]);
            if (r2) return r2;
        }
        return MF_OK;
    };
    double* U = c->d_buf[MF_BUF_U];
    if ((rc = limit(c->d_S))) return rc;
    StageOut s2;
    s2.kind = ST_TWO;
    s2.dt = dt;
    s2.uprev = U;
    s2.out = c->d_T;
    s2.sums_in = s2.sums_out = sums;
    s2.out_dt_off = dt / 2.0;
    if ((rc = run_rhs(c, prec, c->d_S, c->d_buf[MF_BUF_DU], dt, s2))) return rc;
    if ((rc = limit(c->d_T))) return rc;
    StageOut s3;
    s3.kind = ST_THREE;
    s3.dt = dt;
    s3.uprev = U;
    s3.out = U;
    s3.sums_in = s3.sums_out = sums;
    s3.out_dt_off = dt;
    if ((rc = run_rhs(c, prec, c->d_T, c->d_buf[MF_BUF_DU], dt / 2.0, s3))) return rc;
    if ((rc = limit(U))) return rc;
    StageOut s1;
    s1.kind = ST_FSAL;
    s1.dt = dt;
    s1.out = c->d_S;
    s1.sums_in = s1.sums_out = sums;
    s1.out_t_adv = dt;
    s1.out_dt_off = dt;
    if ((rc = run_rhs(c, prec, U, c->d_buf[MF_BUF_DU], dt, s1))) return rc;
    if (history) {
        if ((rc = run_history(c, U, 0, dt))) return rc;  // also advances the time
    } else if ((rc = run_advance_time(c, dt))) {
        return rc;
    }
    return MF_OK;
}

int mf_ssprk33(mf_ctx* c, double t0, double dt, int n_steps, int precision, int flags, int resume)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n_steps < 0 || !(dt > 0.0)) return fail(c, MF_ERR_ARGUMENT, "mf_ssprk33: bad step");
    if ((rc = prepare_rhs(c))) return rc;
    const bool limiter = (flags & MF_STEP_LIMITER) && c->lim_n > 0;
    const int gflags = flags & ~MF_STEP_NOSYNC;  // what the captured step depends on
    const bool history = (flags & MF_STEP_HISTORY) && c->visc_mode == MF_VISC_RESIDUAL;
    const bool sums = stage_sums(c, limiter);
    if (resume) {
        if (!c->fsal_valid || c->fsal_dt != dt)
            return fail(c, MF_ERR_STATE, "mf_ssprk33: resume needs the previous call's state and the same dt");
        // with the residual viscosity the stored FSAL stage carries the
        // next boundary pass and sums exactly when no stage limiter runs
        if (c->fsal_sums != sums)
            return fail(c, MF_ERR_STATE,
                        "mf_ssprk33: resume needs the previous call's stage limiter and viscosity settings");
    } else {
        if ((rc = run_set_time(c, t0, 0))) return rc;
        if (history && (rc = run_history(c, c->d_buf[MF_BUF_U], 1))) return rc;
        StageOut s1;
        s1.kind = ST_FSAL;
        s1.dt = dt;
        s1.out = c->d_S;
        s1.sums_out = sums;  // for the first step's rhs! at t0 + dt
        s1.out_dt_off = dt;
        if ((rc = run_rhs(c, precision, c->d_buf[MF_BUF_U], c->d_buf[MF_BUF_DU], 0.0, s1))) return rc;
        c->host_success_iter = 0;
    }
    if (flags & MF_STEP_GRAPH) {
        // every step reads its time from device memory, so a graph of
        // several consecutive steps replays like the same steps one by one
        auto capture = [&](int steps, hipGraph_t* graph, hipGraphExec_t* exec) -> int {
            MF_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int r = MF_OK;
            for (int s = 0; s < steps && r == MF_OK; ++s) r = ssprk_step(c, dt, precision, limiter, history, sums);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (r) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
            *graph = g;
            MF_HIP(c, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
            return MF_OK;
        };
        if (!c->step_exec || c->graph_prec != precision || c->graph_flags != gflags || c->graph_dt != dt ||
            c->graph_sums != sums) {
            free_graph(c);
            MF_HIP(c, hipStreamSynchronize(c->stream));
            if ((rc = capture(1, &c->step_graph, &c->step_exec))) return rc;
            c->graph_prec = precision;
            c->graph_flags = gflags;
            c->graph_dt = dt;
            c->graph_sums = sums;
        }
        if (n_steps >= kGraphSteps && !c->multi_exec &&
            (rc = capture(kGraphSteps, &c->multi_graph, &c->multi_exec)))
            return rc;
        int n = 0;
        for (; n + kGraphSteps <= n_steps; n += kGraphSteps) MF_HIP(c, hipGraphLaunch(c->multi_exec, c->stream));
        for (; n < n_steps; ++n) MF_HIP(c, hipGraphLaunch(c->step_exec, c->stream));
    } else {
        for (int n = 0; n < n_steps; ++n)
            if ((rc = ssprk_step(c, dt, precision, limiter, history, sums))) return rc;
    }
    c->fsal_dt = dt;
    c->fsal_sums = sums;
    c->host_success_iter += n_steps;
    c->fsal_valid = true;
    if (!(flags & MF_STEP_NOSYNC)) MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

int mf_set_tuning(mf_ctx* c, int key, int precision, int value)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (precision < -1 || precision > MF_PREC_F64_REF)
        return fail(c, MF_ERR_ARGUMENT, "mf_set_tuning: unknown precision");
    const int p0 = precision < 0 ? 0 : precision, p1 = precision < 0 ? MF_PREC_F64_REF : precision;
    if (key == MF_TUNE_FLUX_BATCH) {
        if (value != 0 && value != 3 && value != 5 && value != 6 && value != 10)
            return fail(c, MF_ERR_ARGUMENT, "mf_set_tuning: flux batch must be 0, 3, 5, 6 or 10");
        for (int p = p0; p <= p1; ++p) c->tune_flux_batch[p] = value;
    } else if (key == MF_TUNE_TILES) {
        if (value != 0 && value != 1) return fail(c, MF_ERR_ARGUMENT, "mf_set_tuning: tiles must be 0 or 1");
        MF_HIP(c, hipStreamSynchronize(c->stream));
        c->use_tiles = value;
        free_tiles(c);
        free_graph(c);
        return MF_OK;
    } else if (key == MF_TUNE_NODE_ORDER) {
        if (value != 0 && value != 1) return fail(c, MF_ERR_ARGUMENT, "mf_set_tuning: node order must be 0 or 1");
        c->node_order = value;
        return MF_OK;
    } else if (key == MF_TUNE_WEIGHTS_SOLVER) {
        if (value < 0 || value > 2) return fail(c, MF_ERR_ARGUMENT, "mf_set_tuning: weights solver must be 0, 1 or 2");
        c->weights_solver = value;
        return MF_OK;
    } else if (key == MF_TUNE_TRANSPOSE_BATCH) {
        if (value != 0 && value != 2 && value != 4 && value != 8)
            return fail(c, MF_ERR_ARGUMENT, "mf_set_tuning: transpose batch must be 0, 2, 4 or 8");
        for (int p = p0; p <= p1; ++p) c->tune_transpose_batch[p] = value;
    } else {
        return fail(c, MF_ERR_ARGUMENT, "mf_set_tuning: unknown key");
    }
    free_graph(c);
    return MF_OK;
}

int mf_tile_info(mf_ctx* c, int64_t* info)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!info) return fail(c, MF_ERR_ARGUMENT, "mf_tile_info: null output");
    info[0] = c->tiles_a;
    info[1] = c->tiles_b;
    info[2] = c->n_tiles;
    info[3] = c->tiles_a ? c->ha_total : 0;
    info[4] = c->tiles_b ? c->hb_total : 0;
    info[5] = c->tiles_a ? c->ha_max : 0;
    info[6] = c->tiles_b ? c->hb_max : 0;
    info[7] = c->tiles_b ? c->nb_entries : 0;
    return MF_OK;
}

int mf_time_rhs(mf_ctx* c, double t, int precision, int n_calls, double* rhs_ms, double* kernel_ms)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n_calls < 1) return fail(c, MF_ERR_ARGUMENT, "mf_time_rhs: n_calls < 1");
    if ((rc = prepare_rhs(c))) return rc;
    if ((rc = run_set_time(c, t, -1))) return rc;
    StageOut none;
    // warm, eager
    if ((rc = run_rhs(c, precision, c->d_buf[MF_BUF_U], c->d_buf[MF_BUF_DU], 0.0, none))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    // chains of `reps` back-to-back calls captured in one graph, replayed
    const int reps = std::min(n_calls, 20);
    const int launches = (n_calls + reps - 1) / reps;
    auto time_parts = [&](int parts, double* out_ms) -> int {
        MF_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int r2 = MF_OK;
        for (int q = 0; q < reps && r2 == MF_OK; ++q)
            r2 = run_rhs_parts(c, precision, c->d_buf[MF_BUF_U], c->d_buf[MF_BUF_DU], 0.0, none, parts);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r2 != MF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r2;
        }
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
        hipEvent_t e0, e1;
        MF_HIP(c, hipEventCreate(&e0));
        MF_HIP(c, hipEventCreate(&e1));
        MF_HIP(c, hipGraphLaunch(ex, c->stream));  // warm
        MF_HIP(c, hipEventRecord(e0, c->stream));
        for (int l = 0; l < launches; ++l) MF_HIP(c, hipGraphLaunch(ex, c->stream));
        MF_HIP(c, hipEventRecord(e1, c->stream));
        MF_HIP(c, hipEventSynchronize(e1));
        float ms = 0.f;
        MF_HIP(c, hipEventElapsedTime(&ms, e0, e1));
        *out_ms = ms / ((double)reps * launches);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipGraphExecDestroy(ex);
        return MF_OK;
    };
    double full = 0.0;
    if ((rc = time_parts(RHS_ALL, &full))) return rc;
    if (rhs_ms) *rhs_ms = full;
    // each kernel's own duration inside n_calls further whole rhs! calls:
    // the dispatches carry start/stop events (see launch_k), so every kernel
    // sees the caches as a full call leaves them; batches of 32 calls
    if (kernel_ms) {
        constexpr int kB = 32, kS = 6, kE = 2 * kS;
        std::vector<hipEvent_t> ev(kB * kE, nullptr);
        for (auto& e : ev) MF_HIP(c, hipEventCreate(&e));
        double acc[4] = {0, 0, 0, 0};
        unsigned used[kB];
        int done = 0;
        while (done < n_calls && rc == MF_OK) {
            const int nb = std::min(kB, n_calls - done);
            for (int q = 0; q < nb && rc == MF_OK; ++q) {
                c->probe = &ev[q * kE];
                c->probe_used = 0;
                rc = run_rhs(c, precision, c->d_buf[MF_BUF_U], c->d_buf[MF_BUF_DU], 0.0, none);
                used[q] = c->probe_used;
            }
            c->probe = nullptr;
            if (rc) break;
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) { rc = hip_fail(c, e, "hipStreamSynchronize"); break; }
            for (int q = 0; q < nb; ++q) {
                float m[kS];
                for (int sl = 0; sl < kS; ++sl) {
                    m[sl] = 0.f;
                    if (!(used[q] & (1u << sl))) continue;  // no kernel in this slot
                    e = hipEventElapsedTime(&m[sl], ev[q * kE + 2 * sl], ev[q * kE + 2 * sl + 1]);
                    if (e != hipSuccess) m[sl] = 0.f;
                }
                acc[0] += m[0];         // stencil kernel
                acc[1] += m[1];         // transposed gather
                acc[2] += m[2] + m[3];  // reductions
                acc[3] += m[4] + m[5];  // boundary passes
            }
            done += nb;
        }
        c->probe = nullptr;
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (rc) return rc;
        for (int q = 0; q < 4; ++q) kernel_ms[q] = acc[q] / n_calls;
    }
    c->fsal_valid = false;
    return MF_OK;
}

int mf_time_rhs_cold(mf_ctx* c, double t, int precision, int n_calls, double* rhs_ms)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n_calls < 1 || !rhs_ms) return fail(c, MF_ERR_ARGUMENT, "mf_time_rhs_cold: n_calls < 1 or no output");
    if ((rc = prepare_rhs(c))) return rc;
    if ((rc = run_set_time(c, t, -1))) return rc;
    if (!c->d_flush && (rc = dalloc(c, &c->d_flush, kFlushDoubles))) return rc;
    StageOut none;
    if ((rc = run_rhs(c, precision, c->d_buf[MF_BUF_U], c->d_buf[MF_BUF_DU], 0.0, none))) return rc;
    hipEvent_t e0, e1;
    MF_HIP(c, hipEventCreate(&e0));
    MF_HIP(c, hipEventCreate(&e1));
    std::vector<float> ms;
    for (int q = 0; q < n_calls && rc == MF_OK; ++q) {
        // the flush runs long enough (about 0.2 ms) for the host to enqueue
        // the whole call behind it, so e0..e1 holds device time only
        if ((rc = launch_flush(c, c->d_flush, kFlushDoubles))) break;
        if (hipEventRecord(e0, c->stream) != hipSuccess) {
            rc = fail(c, MF_ERR_HIP, "hipEventRecord");
            break;
        }
        if ((rc = run_rhs(c, precision, c->d_buf[MF_BUF_U], c->d_buf[MF_BUF_DU], 0.0, none))) break;
        float v = 0.f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&v, e0, e1) != hipSuccess) {
            rc = fail(c, MF_ERR_HIP, "timing events");
            break;
        }
        ms.push_back(v);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    std::sort(ms.begin(), ms.end());
    const size_t h = ms.size() / 2;
    *rhs_ms = (ms.size() % 2) ? ms[h] : 0.5 * ((double)ms[h - 1] + (double)ms[h]);
    c->fsal_valid = false;
    return MF_OK;
}

}  // extern "C"
