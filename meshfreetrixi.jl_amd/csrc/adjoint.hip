// Adjoint of rhs! and of SSPRK33: vector-Jacobian products by a
// deterministic gather over the transposed stencil operator (no atomics)
// and the per-node flux Jacobian transpose.  The reference has no adjoint;
// this follows the forward map of rhs.hip exactly:
//   rhs(u) = M ( -sum_d D_d F_d(P u) ),   P = first boundary pass on u,
//   M = second boundary pass on du,
// so  (d rhs/d u)^T lam = P^T sum_d -J_d(P u)^T D_d^T (M lam).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mf_device.h"

namespace mf {

int bc_pass_homog(mf_ctx* c, double* u);
int bc_pass_state(mf_ctx* c, double* u, double dt_off);
int ensure_state(mf_ctx* c);
int run_rhs(mf_ctx* c, int prec, double* d_u, double* d_du, double dt_off, const StageOut& st);

struct VjpArgs {
    int64_t N;
    EqParams eq;
    const int64_t* sell_off;
    const int32_t* sell_w;
    const int32_t* trow;
    const double2* tw;
    const uint8_t* flag;
    const int32_t* entry;
    const double* bnormal;
    const double* x;    // state at which J is evaluated (boundary conditions applied)
    const double* lam;  // cotangent of du, scaled by lam_scale
    double lam_scale;
    const double* add[3];  // out = P^T (sum_q add_scale[q] add[q] + J^T ...)
    double add_scale[3];
    double* out;
    // LDS tiles of the transposed operator (the layout of the rhs! transposed pass)
    int64_t n_tiles;
    const int64_t* hb_off;
    const int32_t* hb_ids;
    const int64_t* sb_off;
    const int32_t* sb_w;
    const uint8_t* rowb;
    const uint16_t* lb;
    const double2* twb;
};

// M^T lam at node r (the second boundary pass is its own transpose):
// Dirichlet rows vanish, slip nodes drop the momentum.
template <int V>
__device__ __forceinline__ void mask_lam(const VjpArgs& a, int64_t r, double* l)
{
    if (!a.flag) return;
    const uint8_t f = a.flag[r];
    if (f & 1u)
#pragma unroll
        for (int v = 0; v < V; ++v) l[v] = 0.0;
    if constexpr (V == 4) {
        if (f & 2u) l[1] = l[2] = 0.0;
    }
}

// J_d(u)^T g for the Euler fluxes (d = 0: x, 1: y), conservative variables.
__device__ __forceinline__ void euler_jt(double gamma, const double* u, const double* gx,
                                         const double* gy, double* o)
{
    const double rho = u[0], m1 = u[1], m2 = u[2], E = u[3];
    const double v1 = m1 / rho, v2 = m2 / rho;
    const double gm = gamma - 1.0;
    const double q = 0.5 * (v1 * v1 + v2 * v2);
    const double p = gm * (E - rho * q);
    const double H = (E + p) / rho;
    // x-flux Jacobian rows
    const double a10 = gm * q - v1 * v1, a11 = (3.0 - gamma) * v1, a12 = -gm * v2, a13 = gm;
    const double a20 = -v1 * v2, a21 = v2, a22 = v1;
    const double a30 = v1 * (gm * q - H), a31 = H - gm * v1 * v1, a32 = -gm * v1 * v2, a33 = gamma * v1;
    // y-flux Jacobian rows
    const double b10 = -v1 * v2, b11 = v2, b12 = v1;
    const double b20 = gm * q - v2 * v2, b21 = -gm * v1, b22 = (3.0 - gamma) * v2, b23 = gm;
    const double b30 = v2 * (gm * q - H), b31 = -gm * v1 * v2, b32 = H - gm * v2 * v2, b33 = gamma * v2;
    o[0] = a10 * gx[1] + a20 * gx[2] + a30 * gx[3] + b10 * gy[1] + b20 * gy[2] + b30 * gy[3];
    o[1] = gx[0] + a11 * gx[1] + a21 * gx[2] + a31 * gx[3] + b11 * gy[1] + b21 * gy[2] + b31 * gy[3];
    o[2] = a12 * gx[1] + a22 * gx[2] + a32 * gx[3] + gy[0] + b12 * gy[1] + b22 * gy[2] + b32 * gy[3];
    o[3] = a13 * gx[1] + a33 * gx[3] + b23 * gy[2] + b33 * gy[3];
}

// Epilogue of one node's product: scale, flux Jacobian transpose at the
// (boundary-imposed) state, the added terms, then P^T.
template <int V>
__device__ __forceinline__ void vjp_finish(const VjpArgs& a, int64_t i, double* gx, double* gy)
{
#pragma unroll
    for (int v = 0; v < V; ++v) {
        gx[v] *= a.lam_scale;
        gy[v] *= a.lam_scale;
    }
    double o[V];
    if constexpr (V == 1) {
        o[0] = -(a.eq.a1 * gx[0] + a.eq.a2 * gy[0]);
    } else {
        double ui[4], jt[4];
        load_state<4>(a.x, i, ui);
        euler_jt(a.eq.gamma, ui, gx, gy, jt);
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = -jt[v];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (!a.add[q]) continue;
        double s[V];
        load_state<V>(a.add[q], i, s);
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = fma(a.add_scale[q], s[v], o[v]);
    }
    // P^T: Dirichlet rows vanish, slip nodes project the momentum
    if (a.flag) {
        const uint8_t f = a.flag[i];
        if (f & 1u)
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = 0.0;
        if constexpr (V == 4) {
            if (f & 2u) {
                const double* nr = a.bnormal + 2 * a.entry[i];
                const double nn = sqrt(nr[0] * nr[0] + nr[1] * nr[1]);
                const double e1 = nr[0] / nn, e2 = nr[1] / nn;
                const double dot = o[1] * e1 + o[2] * e2;
                o[1] -= dot * e1;
                o[2] -= dot * e2;
            }
        }
    }
    store_state<V>(a.out, i, o);
}

template <int V>
__global__ __launch_bounds__(kBlock) void k_vjp(const VjpArgs a)
{
    const int64_t i = remapped_node();
    if (i >= a.N) return;
    const int64_t slice = i >> 6;
    const int lane = (int)(i & 63);
    const int64_t off = a.sell_off[slice];
    const int W = a.sell_w[slice];
    double gx[V], gy[V];
#pragma unroll
    for (int v = 0; v < V; ++v) gx[v] = gy[v] = 0.0;
    for (int e = 0; e < W; ++e) {
        const int64_t q = off + (int64_t)e * kSellC + lane;
        const int32_t r = a.trow[q];
        const double2 w = a.tw[q];
        double l[V];
        load_state<V>(a.lam, r, l);
        mask_lam<V>(a, r, l);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            gx[v] = fma(w.x, l[v], gx[v]);
            gy[v] = fma(w.y, l[v], gy[v]);
        }
    }
    vjp_finish<V>(a, i, gx, gy);
}

// The same product over the LDS tiles of the transposed operator: the
// tile's halo of M^T lam is staged once, the rows walk their entries in
// batches of BT (every load of a batch issued first), in the same order as
// k_vjp, so both give the same bits.
template <int V, int BT>
__global__ __launch_bounds__(kBlock) void k_vjp_tl(const VjpArgs a)
{
    extern __shared__ double lds[];
    const int64_t t = remapped_block();
    if (t >= a.n_tiles) return;  // whole block
    const int64_t h0 = a.hb_off[t];
    const int nh = (int)(a.hb_off[t + 1] - h0);
    for (int q = threadIdx.x; q < nh; q += kBlock) {
        const int64_t r = a.hb_ids[h0 + q];
        double l[V];
        load_state<V>(a.lam, r, l);
        mask_lam<V>(a, r, l);
        if constexpr (V == 4)
            lds_store_unit<1>(lds, q, 0, l);
        else
            store_state<V>(lds, q, l);
    }
    __syncthreads();
    const int pos = threadIdx.x;
    const int64_t i = t * kBlock + a.rowb[t * kBlock + pos];
    if (i >= a.N) return;
    const int lane = pos & 63;
    const int64_t sl = t * (kBlock / kSellC) + (pos >> 6);
    const int64_t off = a.sb_off[sl];
    const int W = a.sb_w[sl];
    double gx[V], gy[V];
#pragma unroll
    for (int v = 0; v < V; ++v) gx[v] = gy[v] = 0.0;
#pragma unroll 1
    for (int e0 = 0; e0 < W; e0 += BT) {
        int32_t r[BT];
        double2 w[BT];
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int e = (e0 + b < W) ? e0 + b : e0;
            const int64_t q = off + (int64_t)e * kSellC + lane;
            r[b] = a.lb[q];
            w[b] = a.twb[q];
        }
        double l[BT][V];
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            if constexpr (V == 4)
                lds_load_unit<1>(lds, r[b], 0, l[b]);
            else
                load_state<V>(lds, r[b], l[b]);
        }
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            if (e0 + b >= W) continue;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                gx[v] = fma(w[b].x, l[b][v], gx[v]);
                gy[v] = fma(w[b].y, l[b][v], gy[v]);
            }
        }
    }
    vjp_finish<V>(a, i, gx, gy);
}

static VjpArgs vjp_args(mf_ctx* c)
{
    VjpArgs a{};
    a.N = c->N;
    a.eq.a1 = c->p0;
    a.eq.a2 = c->p1;
    a.eq.gamma = c->p0;
    a.sell_off = c->d_sell_off;
    a.sell_w = c->d_sell_w;
    a.trow = c->d_trow;
    a.tw = c->d_tw;
    a.flag = c->have_bc ? c->bc.flag : nullptr;
    a.entry = c->have_bc ? c->bc.entry : nullptr;
    a.bnormal = c->have_bc ? c->bc.bnormal : nullptr;
    a.lam_scale = 1.0;
    a.n_tiles = c->n_tiles;
    a.hb_off = c->d_hb_off;
    a.hb_ids = c->d_hb_ids;
    a.sb_off = c->d_sb_off;
    a.sb_w = c->d_sb_w;
    a.rowb = c->d_rowb;
    a.lb = c->d_lb;
    a.twb = c->d_twb;
    return a;
}

static int launch_vjp(mf_ctx* c, const VjpArgs& a)
{
    const unsigned grid = xcd_grid((c->N + kBlock - 1) / kBlock);
    if (c->tiles_b) {
        // halo records of nvar doubles (the tiles were sized for at least that)
        const size_t lds = (size_t)c->hb_max * c->nvar * sizeof(double);
        if (c->nvar == 4)
            hipLaunchKernelGGL((k_vjp_tl<4, 8>), dim3(grid), dim3(kBlock), lds, c->stream, a);
        else
            hipLaunchKernelGGL((k_vjp_tl<1, 8>), dim3(grid), dim3(kBlock), lds, c->stream, a);
    } else if (c->nvar == 4) {
        hipLaunchKernelGGL(k_vjp<4>, dim3(grid), dim3(kBlock), 0, c->stream, a);
    } else {
        hipLaunchKernelGGL(k_vjp<1>, dim3(grid), dim3(kBlock), 0, c->stream, a);
    }
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// L = sum_i rho_i^2 over the (boundary-imposed) state and its gradient
// dL/du = (2 rho, 0, ...).
__global__ void k_loss(int64_t N, int V, const double* u, double* grad, double* loss)
{
    __shared__ double red[1024];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        const double r = u[i * V];
        acc = fma(r, r, acc);
        grad[i * V] = 2.0 * r;
        for (int v = 1; v < V; ++v) grad[i * V + v] = 0.0;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0];
}

static int copy_state(mf_ctx* c, double* dst, const double* src)
{
    MF_HIP(c, hipMemcpyAsync(dst, src, sizeof(double) * c->N * c->nvar, hipMemcpyDeviceToDevice, c->stream));
    return MF_OK;
}

}  // namespace mf

using namespace mf;

int ensure_transpose_any(mf_ctx* c)
{
    if (c->have_T) return MF_OK;
    if (!c->have_flux) return fail(c, MF_ERR_STATE, "no flux operator");
    return launch_build_transpose(c);
}

// State, transposed operator and, where they fit, the LDS tiles of both
// phases (the transposed tiles carry nvar doubles of lam per halo node).
static int prepare_vjp(mf_ctx* c)
{
    int rc;
    if ((rc = ensure_state(c))) return rc;
    if ((rc = ensure_transpose_any(c))) return rc;
    if (c->use_tiles && c->have_flux && (!c->tried_a || (c->tiles_a && !c->tried_b)))
        if ((rc = build_tiles(c, true, c->nvar))) return rc;
    if ((rc = ensure_tile_ops(c))) return rc;
    return MF_OK;
}

extern "C" {

int mf_rhs_vjp(mf_ctx* c, double t)
{
    if (!c) return MF_ERR_ARGUMENT;
    c->err.clear();
    MF_HIP(c, hipSetDevice(c->device));
    int rc;
    if (c->visc_mode != MF_VISC_NONE || c->hv_gamma != 0.0 || c->tom_gamma != 0.0 || c->igr_alpha != 0.0)
        return fail(c, MF_ERR_UNSUPPORTED, "mf_rhs_vjp: sources have no adjoint in this build");
    if ((rc = prepare_vjp(c))) return rc;
    if ((rc = run_set_time(c, t, -1))) return rc;
    // J is evaluated at P u: boundary pass on a copy of u
    if ((rc = copy_state(c, c->d_S, c->d_buf[MF_BUF_U]))) return rc;
    if ((rc = bc_pass_state(c, c->d_S, 0.0))) return rc;
    VjpArgs a = vjp_args(c);
    a.x = c->d_S;
    a.lam = c->d_buf[MF_BUF_AUX0];
    a.out = c->d_buf[MF_BUF_AUX1];
    if ((rc = launch_vjp(c, a))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

// Reverse sweep of SSPRK33 in Shu-Osher form (see rhs.hip):
//   S = U' + dt k0, T = (3U' + S' + dt k1)/4, U+ = (U' + 2T' + 2dt k2)/3
// with U' = P U etc. and k = M f(P .).  Per step, given Ubar+ (of U+):
//   Sbar2 = P^T(2/3 Ubar+ + J(T')^T M (2dt/3 Ubar+))
//   Sbar1 = P^T(1/4 Sbar2 + J(S')^T M (dt/4 Sbar2))
//   Ubar  = P^T(1/3 Ubar+ + 3/4 Sbar2 + Sbar1 + J(U')^T M (dt Sbar1))
int mf_ssprk33_vjp(mf_ctx* c, double t0, double dt, int n_steps, int checkpoint_every, double* loss)
{
    if (!c) return MF_ERR_ARGUMENT;
    c->err.clear();
    MF_HIP(c, hipSetDevice(c->device));
    int rc;
    if (n_steps < 1 || checkpoint_every < 1 || !(dt > 0.0))
        return fail(c, MF_ERR_ARGUMENT, "mf_ssprk33_vjp: bad arguments");
    if (c->visc_mode != MF_VISC_NONE || c->hv_gamma != 0.0 || c->tom_gamma != 0.0 || c->igr_alpha != 0.0)
        return fail(c, MF_ERR_UNSUPPORTED, "mf_ssprk33_vjp: sources have no adjoint in this build");
    if ((rc = prepare_vjp(c))) return rc;
    const int C = std::min(checkpoint_every, n_steps);
    const int nck = (n_steps + C - 1) / C;
    const int64_t NV = c->N * c->nvar;
    // buffers: checkpoints, 3 tape states per step of a segment, Ubar, Sbar2, Sbar1, two work states
    const int64_t nbuf = nck + 3 * (int64_t)C + 6;
    if (c->adj_len < nbuf * NV) {
        dfree(c->d_adj);
        c->d_adj = nullptr;
        c->adj_len = 0;
        if ((rc = dalloc(c, &c->d_adj, (size_t)(nbuf * NV)))) return rc;
        c->adj_len = nbuf * NV;
    }
    // checkpoint s at buf(s); the tape and work states after them
    auto buf = [&](int64_t q) { return c->d_adj + q * NV; };
    double* tape = buf(nck);
    double* Ub = buf(nck + 3 * C);
    double* Ub2 = buf(nck + 3 * C + 1);
    double* Sb2 = buf(nck + 3 * C + 2);
    double* Sb1 = buf(nck + 3 * C + 3);
    double* W0 = buf(nck + 3 * C + 4);
    double* W1 = buf(nck + 3 * C + 5);
    std::vector<double> tn(n_steps + 1);
    tn[0] = t0;
    for (int n = 0; n < n_steps; ++n) tn[n + 1] = tn[n] + dt;
    // one forward step from raw state X (mutated to X'), stages into S and T
    // (left post-boundary), result (raw) into Y
    auto fwd = [&](int n, double* X, double* S, double* T, double* Y) -> int {
        int r2;
        if ((r2 = run_set_time(c, tn[n], -1))) return r2;
        StageOut s1;
        s1.kind = ST_FSAL;
        s1.dt = dt;
        s1.out = S;
        if ((r2 = run_rhs(c, MF_PREC_F64, X, c->d_buf[MF_BUF_DU], 0.0, s1))) return r2;
        StageOut s2;
        s2.kind = ST_TWO;
        s2.dt = dt;
        s2.uprev = X;
        s2.out = T;
        if ((r2 = run_rhs(c, MF_PREC_F64, S, c->d_buf[MF_BUF_DU], dt, s2))) return r2;
        StageOut s3;
        s3.kind = ST_THREE;
        s3.dt = dt;
        s3.uprev = X;
        s3.out = Y;
        return run_rhs(c, MF_PREC_F64, T, c->d_buf[MF_BUF_DU], dt / 2.0, s3);
    };
    // tape of segment step n (first step n0): X at 3 (n - n0), S and T after it
    auto tape_x = [&](int n, int n0) { return tape + (int64_t)(3 * (n - n0)) * NV; };
    // forward pass with checkpoints (raw states at steps 0, C, 2C, ...); the
    // last segment keeps its tape, so the reverse sweep does not recompute it
    const int last0 = (nck - 1) * C;
    if ((rc = copy_state(c, W0, c->d_buf[MF_BUF_U]))) return rc;
    for (int n = 0; n < last0; ++n) {
        if (n % C == 0 && (rc = copy_state(c, buf(n / C), W0))) return rc;
        if ((rc = fwd(n, W0, tape, tape + NV, W1))) return rc;
        std::swap(W0, W1);
    }
    if ((rc = copy_state(c, tape_x(last0, last0), W0))) return rc;
    for (int n = last0; n < n_steps; ++n) {
        double* Xn = tape_x(n, last0);
        if ((rc = fwd(n, Xn, Xn + NV, Xn + 2 * NV, n + 1 < n_steps ? tape_x(n + 1, last0) : W1))) return rc;
    }
    // loss on the final state (W1) after its boundary pass (the last FSAL call)
    if ((rc = run_set_time(c, tn[n_steps], -1))) return rc;
    if ((rc = bc_pass_state(c, W1, 0.0))) return rc;
    double* d_loss = W0;
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(1024), 0, c->stream, c->N, c->nvar, W1, Ub, d_loss);
    MF_HIP(c, hipGetLastError());
    double hl = 0.0;
    MF_HIP(c, hipMemcpyAsync(&hl, d_loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if ((rc = bc_pass_homog(c, Ub))) return rc;
    // reverse sweep, one segment at a time
    for (int sg = nck - 1; sg >= 0; --sg) {
        const int n0 = sg * C, n1 = std::min(n_steps, n0 + C);
        if (sg != nck - 1) {  // recompute the segment's tape from its checkpoint
            if ((rc = copy_state(c, tape_x(n0, n0), buf(sg)))) return rc;
            for (int n = n0; n < n1; ++n) {
                double* Xn = tape_x(n, n0);
                if ((rc = fwd(n, Xn, Xn + NV, Xn + 2 * NV, n + 1 < n1 ? tape_x(n + 1, n0) : W1))) return rc;
            }
        }
        for (int n = n1 - 1; n >= n0; --n) {
            double* Xn = tape + (int64_t)(3 * (n - n0)) * NV;
            double* Sn = Xn + NV;
            double* Tn = Xn + 2 * NV;
            VjpArgs a = vjp_args(c);
            a.x = Tn;
            a.lam = Ub;
            a.lam_scale = 2.0 * dt / 3.0;
            a.add[0] = Ub;
            a.add_scale[0] = 2.0 / 3.0;
            a.out = Sb2;
            if ((rc = launch_vjp(c, a))) return rc;
            VjpArgs b = vjp_args(c);
            b.x = Sn;
            b.lam = Sb2;
            b.lam_scale = dt / 4.0;
            b.add[0] = Sb2;
            b.add_scale[0] = 0.25;
            b.out = Sb1;
            if ((rc = launch_vjp(c, b))) return rc;
            VjpArgs d = vjp_args(c);
            d.x = Xn;
            d.lam = Sb1;
            d.lam_scale = dt;
            d.add[0] = Ub;
            d.add_scale[0] = 1.0 / 3.0;
            d.add[1] = Sb2;
            d.add_scale[1] = 0.75;
            d.add[2] = Sb1;
            d.add_scale[2] = 1.0;
            d.out = Ub2;
            if ((rc = launch_vjp(c, d))) return rc;
            std::swap(Ub, Ub2);
        }
    }
    if ((rc = copy_state(c, c->d_buf[MF_BUF_AUX1], Ub))) return rc;
    MF_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = hl;
    c->fsal_valid = false;
    return MF_OK;
}

}  // extern "C"
