// rhs! on the GPU: strong boundary conditions, the fused flux-divergence /
// hyperviscosity / residual-viscosity stencil kernels, SSPRK33 stage
// updates, the history callback and the Zhang-Shu stage limiter.
//
// Reference: Trixi.rhs!, pointcloudsolver/rbfsolver.jl:397-428 (and the
// functions it calls, cited at each kernel).
//
// Layout: node-major state (V doubles per node), ELL stencils slot-major
// (entry s of node i at s*N + i), operator weights (Dx, Dy) interleaved.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "mf_device.h"

namespace mf {

// Slip-wall projection of the momentum (PointCloudBCs.jl:15-21, :87-106).
__device__ __forceinline__ void slip_project(const double* nrm2, double* s)
{
    const double n1 = nrm2[0], n2 = nrm2[1];
    const double nn = sqrt(n1 * n1 + n2 * n2);
    const double e1 = n1 / nn, e2 = n2 / nn;
    const double m1 = s[1], m2 = s[2];
    const double dot = m1 * e1 + m2 * e2;
    s[1] = fma(-dot, e1, m1);
    s[2] = fma(-dot, e2, m2);
}

// ---------------------------------------- boundary-pass arguments
struct BcArgs {
    int e0, e1;
    const int32_t* bidx;
    const int32_t* btag;
    const double* bnormal;
    const double* btable;
    const double* px;
    const double* py;
    const uint8_t* flag;   // per node (fused use)
    const int32_t* entry;  // per node (fused use)
    double gamma;
    const StepScalars* sc;
    double dt_off;
    double t_adv;  // the pass's time is (sc->t + t_adv) + dt_off (t_adv = 0: sc->t + dt_off)
    int homog;
    const int32_t* tag_kind;  // per tag, device tables
    const int32_t* tag_fn;
    const double* tag_params;  // 8 per tag
    const int32_t* ecode;      // per entry: first-pass kind | function << 4 | tag << 8
};

// The first-pass condition of boundary entry e on the state s of its node;
// false if the entry leaves the state unchanged.
template <int V>
__device__ __forceinline__ bool bc_entry(const BcArgs& b, int e, int32_t node, double* s)
{
    const int tag = b.btag[e];
    const int kind = b.tag_kind[tag];
    if (kind == MF_BC_DIRICHLET) {
        if (b.homog) {
#pragma unroll
            for (int v = 0; v < V; ++v) s[v] = 0.0;
        } else if (b.tag_fn[tag] == MF_FN_TABLE) {
#pragma unroll
            for (int v = 0; v < V; ++v) s[v] = b.btable[(int64_t)e * V + v];
        } else {
            double t = b.sc->t;
            if (b.t_adv != 0.0) t = t + b.t_adv;
            t = t + b.dt_off;
            bc_value<V>(b.tag_fn[tag], b.tag_params + 8 * tag, b.gamma, b.px[node], b.py[node], t, s);
        }
        return true;
    }
    if (kind == MF_BC_SLIP_WALL) {
        if constexpr (V == 4) {
            slip_project(b.bnormal + 2 * e, s);
            return true;
        }
    }
    return false;
}

// bc_entry's first pass for node `node` with boundary entry e, from the
// entry's packed code: the code, the node's coordinates and the time load
// together, then the tag's parameters (or the entry's table row or normal):
// two dependent loads after the entry index instead of bc_entry's four;
// same arithmetic.
template <int V>
__device__ __forceinline__ bool bc_node(const BcArgs& b, int e, int32_t node, double* s)
{
    const int32_t code = b.ecode[e];
    const double xn = b.px[node], yn = b.py[node], t0 = b.sc->t;
    const int kind = code & 15, fn = (code >> 4) & 15, tag = code >> 8;
    if (kind == MF_BC_DIRICHLET) {
        if (b.homog) {
#pragma unroll
            for (int v = 0; v < V; ++v) s[v] = 0.0;
        } else if (fn == MF_FN_TABLE) {
#pragma unroll
            for (int v = 0; v < V; ++v) s[v] = b.btable[(int64_t)e * V + v];
        } else {
            double t = t0;
            if (b.t_adv != 0.0) t = t + b.t_adv;
            t = t + b.dt_off;
            bc_value<V>(fn, b.tag_params + 8 * tag, b.gamma, xn, yn, t, s);
        }
        return true;
    }
    if (kind == MF_BC_SLIP_WALL) {
        if constexpr (V == 4) {
            slip_project(b.bnormal + 2 * e, s);
            return true;
        }
    }
    return false;
}

// ----------------------------------------------------------- arguments
struct RhsArgs {
    int64_t N;
    int k;
    const int32_t* __restrict__ idx;
    const int16_t* __restrict__ didx;
    const double2* __restrict__ wxy;
    const double* __restrict__ hv;
    const float2* __restrict__ wxy32;
    const float* __restrict__ hv32;
    double hv_gamma;
    EqParams eq;
    const double* __restrict__ u;
    double* __restrict__ du;
    const uint8_t* __restrict__ flag;
    const int32_t* __restrict__ entry;
    const double* __restrict__ bnormal;
    // viscosity
    int visc;
    int hv_first;
    double c_rv, c_uw, h;
    const double* __restrict__ approx;
    const StepScalars* __restrict__ sc;
    const double* __restrict__ pmax;  // per-block lexicographic maxima of the max pass
    int nparts;
    double* norms_out;  // ninf published here (inspection)
    double* __restrict__ G;
    float* __restrict__ G32;   // split-FP32: G in FP32 (same bytes as the first half of G)
    float* __restrict__ du32;  // split-FP32: phase A's du in FP32
    double* __restrict__ eps;
    int8_t* __restrict__ epsc;
    // column-sorted operators (reference operation order)
    const int32_t* __restrict__ idx_s;
    const double* __restrict__ wx_s;
    const double* __restrict__ wy_s;
    const double* __restrict__ hv_s;
    // transposed operator
    const int64_t* __restrict__ sell_off;
    const int32_t* __restrict__ sell_w;
    const int32_t* __restrict__ trow;
    const double2* __restrict__ tw;
    const float2* __restrict__ tw32;
    // Tominec hyperviscosity: phase A stores L = lap u, phase B gathers lap' L
    const double* __restrict__ lap;
    const float* __restrict__ lap32;
    const double* __restrict__ lap_s;  // column-sorted
    double tom_gamma;
    double* __restrict__ L;
    float* __restrict__ L32;           // split-FP32: L in FP32 (same buffer)
    const double* __restrict__ tl;     // lap of the transposed entries
    const float* __restrict__ tl32;
    const double* __restrict__ twl;    // the same in the tile layout
    const float* __restrict__ twl32;
    // reference order: Flyer term in phase A (it precedes every phase-B
    // source); phase-B sources in keyword order, 2 bits each from the low
    // end (kSeqViscosity, kSeqTominec, kSeqFlyer)
    int flyer_a;
    int seq_b;
    // LDS tiles (phase A: halo of the stencils; phase B: halo of D^T)
    int64_t n_tiles;
    const int64_t* __restrict__ ha_off;
    const int32_t* __restrict__ ha_ids;
    const uint16_t* __restrict__ la;
    const uint16_t* __restrict__ la_s;  // halo positions of the column-sorted stencil
    const int64_t* __restrict__ hb_off;
    const int32_t* __restrict__ hb_ids;
    const int64_t* __restrict__ sb_off;
    const int32_t* __restrict__ sb_w;
    const uint8_t* __restrict__ rowb;
    const uint16_t* __restrict__ lb;
    const double2* __restrict__ twb;
    const float2* __restrict__ twb32;
    // stage output
    int st_kind;
    double dt;
    const double* __restrict__ uprev;
    double* __restrict__ out;
    // stepping: the output is the next rhs!'s input.  With stage_psum set
    // the transposed kernel applies that rhs!'s first boundary pass (obc)
    // to the output and stores the output's per-block sums there (tile or
    // node block t at 4 t), which that rhs!'s max pass reads in place of
    // its own sum pass
    double* __restrict__ stage_psum;
    BcArgs obc;
};

// FP32 node records (split-FP32 intermediates): V floats per record.
template <int V>
__device__ __forceinline__ void store_state32(float* __restrict__ p, int64_t j, const double* s)
{
    if constexpr (V == 4) {
        reinterpret_cast<float4*>(p)[j] = make_float4((float)s[0], (float)s[1], (float)s[2], (float)s[3]);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) p[j * V + v] = (float)s[v];
    }
}

template <int V>
__device__ __forceinline__ void load_state32(const float* __restrict__ p, int64_t j, double* s)
{
    if constexpr (V == 4) {
        const float4 q = reinterpret_cast<const float4*>(p)[j];
        s[0] = q.x;
        s[1] = q.y;
        s[2] = q.z;
        s[3] = q.w;
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) s[v] = p[j * V + v];
    }
}

// Second boundary pass on du (rbfsolver.jl:422-425): Dirichlet nodes get
// du = 0, slip nodes du momentum = 0.
template <int V>
__device__ __forceinline__ void mask_du(const RhsArgs& a, int64_t i, double* du)
{
    if (!a.flag) return;
    const uint8_t f = a.flag[i];
    if (f & 1u) {
#pragma unroll
        for (int v = 0; v < V; ++v) du[v] = 0.0;
    }
    if constexpr (V == 4) {
        if (f & 2u) {
            du[1] = 0.0;
            du[2] = 0.0;
        }
    }
}

// Writes du, or the next SSPRK33 stage built from k = du at this node:
//   FSAL : out = u + dt k
//   TWO  : out = (3 uprev + u + dt k) / 4
//   THREE: out = (uprev + 2 u + 2 dt k) / 3
// u is this node's (boundary-imposed) input state; slip nodes see the
// second boundary pass's projection first, as the reference's u does.
template <int V>
__device__ __forceinline__ void stage_value(const RhsArgs& a, int64_t i, const double* du, double* o)
{
    double ui[V];
    load_state<V>(a.u, i, ui);
    if constexpr (V == 4) {
        if (a.flag && (a.flag[i] & 2u)) slip_project(a.bnormal + 2 * a.entry[i], ui);
    }
    if (a.st_kind == ST_FSAL) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = fma(a.dt, du[v], ui[v]);
    } else {
        double up[V];
        load_state<V>(a.uprev, i, up);
        if (a.st_kind == ST_TWO) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = fma(3.0, up[v], fma(a.dt, du[v], ui[v])) / 4.0;
        } else {
            const double dt2 = 2.0 * a.dt;
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = fma(2.0, ui[v], fma(dt2, du[v], up[v])) / 3.0;
        }
    }
}

template <int V>
__device__ __forceinline__ void emit(const RhsArgs& a, int64_t i, const double* du)
{
    if (a.st_kind == ST_NONE) {
        store_state<V>(a.du, i, du);
        return;
    }
    double o[V];
    stage_value<V>(a, i, du, o);
    store_state<V>(a.out, i, o);
}

// emit() of a stage whose output the next rhs! reads (a.stage_psum set):
// that rhs!'s first boundary pass is applied to the output here, as its
// sum pass would apply it, and o returns the stored state for the sums.
__device__ __forceinline__ void emit_staged(const RhsArgs& a, int64_t i, const double* du, double* o)
{
    stage_value<4>(a, i, du, o);
    if (a.obc.flag) {
        const int32_t e = a.obc.entry[i];
        if (e >= 0) bc_node<4>(a.obc, e, (int32_t)i, o);
    }
    store_state<4>(a.out, i, o);
}

// Residual/upwind viscosity at node i (hyperviscosity.jl:246-349): eps_uw
// from the local state, eps_rv from |approx_du - du| scaled by the global
// norms, then the choice of update_visc!.
__device__ __forceinline__ double node_viscosity(const RhsArgs& a, int64_t i, const double* ui,
                                                 const double* dui, const double* ninf, int8_t* code)
{
    const double gam = a.eq.gamma;
    const double rho = ui[0], m1 = ui[1], m2 = ui[2], E = ui[3];
    const double v1 = m1 / rho, v2 = m2 / rho;
    double p = (gam - 1.0) * fma(-0.5, fma(m1, v1, m2 * v2), E);
    const double speed = sqrt(v1 * v1 + v2 * v2);
    double sound = 0.0;
    if (!(p < 0.0 || rho < 0.0)) sound = sqrt(gam * p / rho);
    const double eps_uw = a.c_uw * 0.5 * a.h * (speed + sound);
    if (a.visc == MF_VISC_UPWIND) {
        *code = 1;
        return eps_uw;
    }
    double ad[4];
    load_state<4>(a.approx, i, ad);
    double mr = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double r = fabs(ad[v] - dui[v]) / ninf[v];
        mr = (v == 0) ? r : jmax(mr, r);
    }
    const double eps_rv = 0.5 * a.c_rv * (a.h * a.h) * mr;
    if (isnan(eps_rv) || isinf(eps_rv) || a.sc->hist_si == 0) {
        if (isnan(eps_uw) || isinf(eps_uw)) {
            *code = 2;
            return DBL_EPSILON;
        }
        *code = 1;
        return eps_uw;
    }
    *code = eps_rv < eps_uw ? 0 : 1;
    return eps_rv < eps_uw ? eps_rv : eps_uw;
}

constexpr int kSeqViscosity = 1, kSeqTominec = 2, kSeqFlyer = 3;

// ------------------------------------------------------------ FP64 kernel
// du_i = -sum_j Dx_ij F1(u_j) - sum_j Dy_ij F2(u_j)     (calc_fluxes!, :247-265)
//        - gamma sum_j H_ij u_j                         (Flyer HV, hyperviscosity.jl:52-64)
// HVK bit 0: Flyer source; bit 1: Tominec source.
// With a viscosity source this is phase A: it also stores
// G_i = eps_i (Dx u, Dy u)_i for the transposed pass (hyperviscosity.jl:401-407).
// With the Tominec source it is phase A too and stores L_i = sum_j lap_ij u_j
// for the transposed pass du_i -= gamma sum_r lap_ri L_r (hyperviscosity.jl:
// 111-134, the factors of H = lap' lap).
// Residual-viscosity norms in the stencil kernels: thread b loads the max
// pass's partial b at kernel start (at most kBlock partials) and keeps it
// in registers untouched, so the load overlaps the halo staging and the
// stencil loop instead of holding them up; the block reduces the partials
// just before the viscosity is needed.  Every block gets the same bits.
__device__ __forceinline__ void prefetch_max_partials(const RhsArgs& a, double* pm)
{
    if ((int)threadIdx.x < a.nparts)
#pragma unroll
        for (int v = 0; v < 4; ++v) pm[v] = a.pmax[4 * threadIdx.x + v];
}

__device__ __forceinline__ void finish_norms(const RhsArgs& a, double* pm, double* ninf)
{
    block_lexmax4(pm);
#pragma unroll
    for (int v = 0; v < 4; ++v) ninf[v] = (pm[v] == 0.0) ? DBL_EPSILON : pm[v];
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.norms_out)
#pragma unroll
        for (int v = 0; v < 4; ++v) a.norms_out[4 + v] = ninf[v];
}


// Halo record of a node in LDS: for the Euler equations (rho, m1, m2, E,
// v1, v2) with the velocity evaluated exactly as flux_xy does, which takes
// the divisions out of the per-entry work; the pressure is formed from the
// record per entry with flux_xy's expression (so the fluxes have the same
// bits).  48 bytes, three 16-byte chunks: a stride that spreads the
// scattered reads of a wavefront over all LDS banks, and a halo of 3/4 the
// size of the 64-byte record with the pressure.  Otherwise the state.
template <int V>
constexpr int rec_len()
{
    return V == 4 ? 6 : V;
}

template <int V>
__device__ __forceinline__ void make_record(const EqParams& e, const double* u, double* r)
{
    if constexpr (V == 4) {
        const double rho = u[0], m1 = u[1], m2 = u[2], E = u[3];
        const double v1 = m1 / rho, v2 = m2 / rho;
        r[0] = rho;
        r[1] = m1;
        r[2] = m2;
        r[3] = E;
        r[4] = v1;
        r[5] = v2;
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) r[v] = u[v];
    }
}

template <int V>
__device__ __forceinline__ void flux_record(const EqParams& e, const double* r, double* f1, double* f2)
{
    if constexpr (V == 4) {
        const double m1 = r[1], m2 = r[2], E = r[3], v1 = r[4], v2 = r[5];
        const double p = (e.gamma - 1.0) * fma(-0.5, fma(m1, v1, m2 * v2), E);
        const double ep = E + p;
        f1[0] = m1;
        f1[1] = fma(m1, v1, p);
        f1[2] = m1 * v2;
        f1[3] = ep * v1;
        f2[0] = m2;
        f2[1] = m2 * v1;
        f2[2] = fma(m2, v2, p);
        f2[3] = ep * v2;
    } else {
        flux_xy<V>(e, r, f1, f2);
    }
}

template <int V>
__device__ __forceinline__ void load_record(const double* lds, int64_t l, double* r)
{
    if constexpr (V == 4) {
        const double2* p = reinterpret_cast<const double2*>(lds) + 3 * l;
        const double2 c0 = p[0], c1 = p[1], c2 = p[2];
        r[0] = c0.x;
        r[1] = c0.y;
        r[2] = c1.x;
        r[3] = c1.y;
        r[4] = c2.x;
        r[5] = c2.y;
    } else {
        load_state<V>(lds, l, r);
    }
}

// Neighbour j of a stencil loop: its halo record (TL) or its state.
template <int V, bool TL>
__device__ __forceinline__ void load_nb(const RhsArgs& a, const double* lds, int64_t j, double* r)
{
    if constexpr (TL)
        load_record<V>(lds, j, r);
    else
        load_state<V>(a.u, j, r);
}

template <int V, bool TL>
__device__ __forceinline__ void flux_nb(const EqParams& e, const double* r, double* f1, double* f2)
{
    if constexpr (TL)
        flux_record<V>(e, r, f1, f2);
    else
        flux_xy<V>(e, r, f1, f2);
}

// Node of this thread and, with LDS tiles (TL), the tile's stencil halo
// staged in lds: halo entry q holds the record of node ha_ids[ha_off[t] + q].
// Returns -1 for threads without a node (after the block barrier) and
// kNoBlock when the whole block has none.
constexpr int64_t kNoBlock = -2;

template <int V, bool TL>
__device__ __forceinline__ int64_t stage_halo_a(const RhsArgs& a, double* lds)
{
    if constexpr (TL) {
        const int64_t t = remapped_block();
        if (t >= a.n_tiles) return kNoBlock;  // whole block
        const int64_t h0 = a.ha_off[t];
        const int nh = (int)(a.ha_off[t + 1] - h0);
        for (int q = threadIdx.x; q < nh; q += kBlock) {
            double s[V], r[rec_len<V>()];
            load_state<V>(a.u, a.ha_ids[h0 + q], s);
            make_record<V>(a.eq, s, r);
            if constexpr (V == 4) {
                double2* p = reinterpret_cast<double2*>(lds) + 3 * q;
                p[0] = make_double2(r[0], r[1]);
                p[1] = make_double2(r[2], r[3]);
                p[2] = make_double2(r[4], r[5]);
            } else {
                store_state<V>(lds, q, r);
            }
        }
        __syncthreads();
        const int64_t i = t * kBlock + threadIdx.x;
        return i < a.N ? i : -1;
    } else {
        if (remapped_block() * kBlock >= a.N) return kNoBlock;  // whole block
        const int64_t i = remapped_node();
        return i < a.N ? i : -1;
    }
}

// Batches of BT stencil entries: every load of a batch (positions,
// weights, then the gathered states) is issued before any of it is used,
// so BT gathers are in flight per thread.  Entries past K read a valid node
// with weight 0.  TL: neighbour records come from the LDS halo, addressed
// by 16-bit halo positions (same slot-major layout as idx).
template <int V, int NS, int HVK, bool RV, int BT, bool TL>
__global__ __launch_bounds__(kBlock) void k_rhs_f64(const RhsArgs a)
{
    constexpr bool FLY = HVK & 1, TOM = HVK & 2;
    extern __shared__ double lds[];
    double pm[4] = {-1.0, -1.0, -1.0, -1.0};
    const bool rvres = RV && a.visc == MF_VISC_RESIDUAL;
    if (rvres) prefetch_max_partials(a, pm);
    const int64_t i_node = stage_halo_a<V, TL>(a, lds);
    if (i_node == kNoBlock) return;  // whole block
    // a thread without a node shadows the last node until the block-wide
    // norm reduction, then leaves
    const bool live = i_node >= 0;
    if (!rvres && !live) return;
    const int64_t i = live ? i_node : a.N - 1;
    const int64_t N = a.N;
    const int K = NS ? NS : a.k;
    double ax[V], ay[V], gx[V], gy[V], hh[V], ll[V];
#pragma unroll
    for (int v = 0; v < V; ++v) ax[v] = ay[v] = gx[v] = gy[v] = hh[v] = ll[v] = 0.0;
#pragma unroll 1
    for (int s0 = 0; s0 < K; s0 += BT) {
        int32_t j[BT];
        double2 w[BT];
        double hw[BT], lw[BT];
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int s = s0 + b;
            if ((NS && NS % BT == 0) || s < K) {
                const int64_t q = TL ? tm_pos(s, i, K) : (int64_t)s * N + i;
                if constexpr (TL)
                    j[b] = a.la[q];
                else
                    j[b] = a.idx[q];
                w[b] = a.wxy[q];
                if constexpr (FLY) hw[b] = a.hv[q];
                if constexpr (TOM) lw[b] = a.lap[q];
            } else {
                j[b] = TL ? 0 : (int32_t)i;
                w[b] = make_double2(0.0, 0.0);
                hw[b] = lw[b] = 0.0;
            }
        }
        double uj[BT][TL ? rec_len<V>() : V];
#pragma unroll
        for (int b = 0; b < BT; ++b) load_nb<V, TL>(a, lds, j[b], uj[b]);
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            double f1[V], f2[V];
            flux_nb<V, TL>(a.eq, uj[b], f1, f2);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                ax[v] = fma(w[b].x, f1[v], ax[v]);
                ay[v] = fma(w[b].y, f2[v], ay[v]);
            }
            if constexpr (RV) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    gx[v] = fma(w[b].x, uj[b][v], gx[v]);
                    gy[v] = fma(w[b].y, uj[b][v], gy[v]);
                }
            }
            if constexpr (FLY) {
#pragma unroll
                for (int v = 0; v < V; ++v) hh[v] = fma(hw[b], uj[b][v], hh[v]);
            }
            if constexpr (TOM) {
#pragma unroll
                for (int v = 0; v < V; ++v) ll[v] = fma(lw[b], uj[b][v], ll[v]);
            }
        }
    }
    double du[V];
#pragma unroll
    for (int v = 0; v < V; ++v) du[v] = -ax[v] - ay[v];
    if constexpr (FLY) {
        if (!RV || a.hv_first) {
#pragma unroll
            for (int v = 0; v < V; ++v) du[v] = fma(-a.hv_gamma, hh[v], du[v]);
        }
    }
    if constexpr (RV || TOM) {
        if constexpr (RV) {
            double ui[V];
            load_state<V>(a.u, i, ui);
            int8_t code = 1;
            double ninf[4] = {1.0, 1.0, 1.0, 1.0};
            if (rvres) finish_norms(a, pm, ninf);
            if (!live) return;
            const double e = node_viscosity(a, i, ui, du, ninf, &code);
            a.eps[i] = e;
            a.epsc[i] = code;
            double g[2 * V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                g[v] = e * gx[v];
                g[V + v] = e * gy[v];
            }
            store_state<V>(a.G, 2 * i, g);
            store_state<V>(a.G, 2 * i + 1, g + V);
            if constexpr (FLY) {
                if (!a.hv_first) {
#pragma unroll
                    for (int v = 0; v < V; ++v) du[v] = fma(-a.hv_gamma, hh[v], du[v]);
                }
            }
        }
        if constexpr (TOM) store_state<V>(a.L, i, ll);
        store_state<V>(a.du, i, du);
        return;
    }
    mask_du<V>(a, i, du);
    emit<V>(a, i, du);
}

// --------------------------------------- FP64 in the reference's order
// The same operator evaluated in the operation order of the reference's
// SparseArrays products (pointcloudsolver/rbfsolver.jl:21-27, :247-265;
// hyperviscosity.jl:61-63, :401-407): for each node, the x-flux terms in
// ascending column order, du = du - w*F with the product rounded first,
// then the y-flux terms onto that result, then (Flyer) the terms
// du = du + h*(u_j*(-gamma)); D_d u accumulated the same way.  Bitwise
// reproducible against a CSC implementation with the same weights.
// One pass over the column-sorted stencil in batches of BT entries (loads
// first, then the adds in entry order).  PASS 0: x flux (and D_x u),
// 1: y flux (and D_y u), 2: Flyer term, 3: L = lap u of the Tominec source
// (lap * u in CSC order, as the first factor of its H = lap' lap).
template <int V, int NS, bool RV, int BT, int PASS, bool TL>
__device__ __forceinline__ void ref_pass(const RhsArgs& a, int64_t i, int K, double* du, double* g,
                                         const double* lds)
{
    const int64_t N = a.N;
    const double* __restrict__ wsrc = PASS == 0 ? a.wx_s : (PASS == 1 ? a.wy_s : (PASS == 2 ? a.hv_s : a.lap_s));
    const double ng = -a.hv_gamma;
#pragma unroll 1
    for (int s0 = 0; s0 < K; s0 += BT) {
        int32_t j[BT];
        double w[BT];
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int s = s0 + b;
            const bool in = (NS && NS % BT == 0) || s < K;
            const int64_t q = TL ? tm_pos(in ? s : s0, i, K) : (int64_t)(in ? s : s0) * N + i;
            if constexpr (TL)
                j[b] = a.la_s[q];
            else
                j[b] = a.idx_s[q];
            w[b] = wsrc[q];
        }
        double uj[BT][TL ? rec_len<V>() : V];
#pragma unroll
        for (int b = 0; b < BT; ++b) load_nb<V, TL>(a, lds, j[b], uj[b]);
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            if (!((NS && NS % BT == 0) || s0 + b < K)) continue;
            if constexpr (PASS == 2) {
#pragma unroll
                for (int v = 0; v < V; ++v) du[v] = du[v] + w[b] * (uj[b][v] * ng);
            } else if constexpr (PASS == 3) {
#pragma unroll
                for (int v = 0; v < V; ++v) du[v] = du[v] + w[b] * uj[b][v];
            } else {
                double f1[V], f2[V];
                flux_nb<V, TL>(a.eq, uj[b], f1, f2);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    du[v] = du[v] - w[b] * (PASS == 0 ? f1[v] : f2[v]);
                    if constexpr (RV) g[v] = g[v] + w[b] * uj[b][v];
                }
            }
        }
    }
}

template <int V, int NS, int HVK, bool RV, int BT, bool TL>
__global__ __launch_bounds__(kBlock) void k_rhs_ref(const RhsArgs a)
{
    constexpr bool FLY = HVK & 1, TOM = HVK & 2;
    extern __shared__ double lds[];
    double pm[4] = {-1.0, -1.0, -1.0, -1.0};
    const bool rvres = RV && a.visc == MF_VISC_RESIDUAL;
    if (rvres) prefetch_max_partials(a, pm);
    const int64_t i_node = stage_halo_a<V, TL>(a, lds);
    if (i_node == kNoBlock) return;  // whole block
    // a thread without a node shadows the last node until the block-wide
    // norm reduction, then leaves
    const bool live = i_node >= 0;
    if (!rvres && !live) return;
    const int64_t i = live ? i_node : a.N - 1;
    const int K = NS ? NS : a.k;
    double du[V], gx[V], gy[V];
#pragma unroll
    for (int v = 0; v < V; ++v) du[v] = gx[v] = gy[v] = 0.0;
    ref_pass<V, NS, RV, BT, 0, TL>(a, i, K, du, gx, lds);
    ref_pass<V, NS, RV, BT, 1, TL>(a, i, K, du, gy, lds);
    if constexpr (FLY) {
        if (a.flyer_a) ref_pass<V, NS, false, BT, 2, TL>(a, i, K, du, nullptr, lds);
    }
    if constexpr (RV || TOM) {
        if constexpr (RV) {
            double ui[V];
            load_state<V>(a.u, i, ui);
            int8_t code = 1;
            double ninf[4] = {1.0, 1.0, 1.0, 1.0};
            if (rvres) finish_norms(a, pm, ninf);
            if (!live) return;
            const double e = node_viscosity(a, i, ui, du, ninf, &code);
            a.eps[i] = e;
            a.epsc[i] = code;
            double g[2 * V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                g[v] = e * gx[v];
                g[V + v] = e * gy[v];
            }
            store_state<V>(a.G, 2 * i, g);
            store_state<V>(a.G, 2 * i + 1, g + V);
        }
        if constexpr (TOM) {
            double ll[V];
#pragma unroll
            for (int v = 0; v < V; ++v) ll[v] = 0.0;
            ref_pass<V, NS, false, BT, 3, TL>(a, i, K, ll, nullptr, lds);
            store_state<V>(a.L, i, ll);
        }
        store_state<V>(a.du, i, du);
        return;
    }
    mask_du<V>(a, i, du);
    emit<V>(a, i, du);
}

// ------------------------------------------------- split-FP32 kernel
// Same operator; the stencil sums run in FP32 on centred differences
// formed in FP64: sum_j w_ij F(u_j) ~= sum_j w_ij (F(u_j) - F(u_i)),
// exact up to F(u_i) sum_j w_ij, which vanishes with the weights' row sums.
// Weights FP32, neighbour indices as int16 offsets when the cloud allows.
template <int V, int NS, int HVK, bool RV, bool D16, int BT, bool TL>
__global__ __launch_bounds__(kBlock) void k_rhs_s32(const RhsArgs a)
{
    constexpr bool FLY = HVK & 1, TOM = HVK & 2;
    extern __shared__ double lds[];
    double pm[4] = {-1.0, -1.0, -1.0, -1.0};
    const bool rvres = RV && a.visc == MF_VISC_RESIDUAL;
    if (rvres) prefetch_max_partials(a, pm);
    const int64_t i_node = stage_halo_a<V, TL>(a, lds);
    if (i_node == kNoBlock) return;  // whole block
    // a thread without a node shadows the last node until the block-wide
    // norm reduction, then leaves
    const bool live = i_node >= 0;
    if (!rvres && !live) return;
    const int64_t i = live ? i_node : a.N - 1;
    const int64_t N = a.N;
    const int K = NS ? NS : a.k;
    double ui[V], fi1[V], fi2[V];
    load_state<V>(a.u, i, ui);
    flux_xy<V>(a.eq, ui, fi1, fi2);
    float ax[V], ay[V], gx[V], gy[V], hh[V], ll[V];
#pragma unroll
    for (int v = 0; v < V; ++v) ax[v] = ay[v] = gx[v] = gy[v] = hh[v] = ll[v] = 0.0f;
#pragma unroll 1
    for (int s0 = 0; s0 < K; s0 += BT) {
        int64_t jj[BT];
        float2 ww[BT];
        float hw32[BT], lw32[BT];
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int s = s0 + b;
            if ((NS && NS % BT == 0) || s < K) {
                const int64_t q = TL ? tm_pos(s, i, K) : (int64_t)s * N + i;
                if constexpr (TL)
                    jj[b] = a.la[q];
                else if constexpr (D16)
                    jj[b] = i + (int64_t)a.didx[q];
                else
                    jj[b] = a.idx[q];
                ww[b] = a.wxy32[q];
                if constexpr (FLY) hw32[b] = a.hv32[q];
                if constexpr (TOM) lw32[b] = a.lap32[q];
            } else {
                jj[b] = TL ? 0 : i;
                ww[b] = make_float2(0.0f, 0.0f);
                hw32[b] = lw32[b] = 0.0f;
            }
        }
        // keep the batch's index and weight loads issued together, ahead of
        // the neighbour reads (the scheduler otherwise sinks each one to its use)
        __builtin_amdgcn_sched_barrier(0);
        double ub[BT][TL ? rec_len<V>() : V];
#pragma unroll
        for (int b = 0; b < BT; ++b) load_nb<V, TL>(a, lds, jj[b], ub[b]);
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const float2 w = ww[b];
            const double* uj = ub[b];
            double f1[V], f2[V];
            flux_nb<V, TL>(a.eq, uj, f1, f2);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                ax[v] = fmaf(w.x, (float)(f1[v] - fi1[v]), ax[v]);
                ay[v] = fmaf(w.y, (float)(f2[v] - fi2[v]), ay[v]);
            }
            if constexpr (RV || FLY || TOM) {
                float dd[V];
#pragma unroll
                for (int v = 0; v < V; ++v) dd[v] = (float)(uj[v] - ui[v]);
                if constexpr (RV) {
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        gx[v] = fmaf(w.x, dd[v], gx[v]);
                        gy[v] = fmaf(w.y, dd[v], gy[v]);
                    }
                }
                if constexpr (FLY) {
                    const float hw = hw32[b];
#pragma unroll
                    for (int v = 0; v < V; ++v) hh[v] = fmaf(hw, dd[v], hh[v]);
                }
                if constexpr (TOM) {
                    const float lw = lw32[b];
#pragma unroll
                    for (int v = 0; v < V; ++v) ll[v] = fmaf(lw, dd[v], ll[v]);
                }
            }
        }
    }
    double du[V];
#pragma unroll
    for (int v = 0; v < V; ++v) du[v] = -(double)ax[v] - (double)ay[v];
    if constexpr (FLY) {
        if (!RV || a.hv_first) {
#pragma unroll
            for (int v = 0; v < V; ++v) du[v] = fma(-a.hv_gamma, (double)hh[v], du[v]);
        }
    }
    if constexpr (RV || TOM) {
        if constexpr (RV) {
            int8_t code = 1;
            double ninf[4] = {1.0, 1.0, 1.0, 1.0};
            if (rvres) finish_norms(a, pm, ninf);
            if (!live) return;
            const double e = node_viscosity(a, i, ui, du, ninf, &code);
            a.eps[i] = e;
            a.epsc[i] = code;
            double g[2 * V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                g[v] = e * (double)gx[v];
                g[V + v] = e * (double)gy[v];
            }
            // split-FP32: G, L and the partial du go to phase B in FP32
            store_state32<V>(a.G32, 2 * i, g);
            store_state32<V>(a.G32, 2 * i + 1, g + V);
            if constexpr (FLY) {
                if (!a.hv_first) {
#pragma unroll
                    for (int v = 0; v < V; ++v) du[v] = fma(-a.hv_gamma, (double)hh[v], du[v]);
                }
            }
        }
        if constexpr (TOM) {
            double l[V];
#pragma unroll
            for (int v = 0; v < V; ++v) l[v] = (double)ll[v];
            store_state32<V>(a.L32, i, l);
        }
        store_state32<V>(a.du32, i, du);
        return;
    }
    mask_du<V>(a, i, du);
    emit<V>(a, i, du);
}

// ------------------------------------------ phase B: transposed gather
// du_i -= sum_r Dx_ri Gx_r ; du_i -= sum_r Dy_ri Gy_r   (the D_d' products of
// hyperviscosity.jl:405; GT bit 0) and du_i -= gamma sum_r lap_ri L_r (the
// Tominec source, hyperviscosity.jl:111-134; GT bit 1), over the sorted
// SELL-64 transpose, then the second boundary pass and the output.
// TM: 0 FP64 (fused multiply-add), 1 FP32 weights, G and L, 2 FP64 in the
// reference's order (tmp = tmp + w*G over ascending rows, then du - tmp;
// the phase-B sources added in keyword order, a Flyer source among them by
// its own gather of u).
// TL: the tile's rows in the order of rowb (by degree), SELL-64 per
// wavefront (sb_off, sb_w), entries addressing halo records in LDS (lb:
// halo positions; record q belongs to node hb_ids[hb_off[t] + q] and holds
// its G (2V values) and / or L (V values)).
template <int V, int TM, int BT, bool TL, int GT>
__global__ __launch_bounds__(kBlock) void k_rhs_transpose(const RhsArgs a)
{
    constexpr bool HG = GT & 1, HL = GT & 2;
    constexpr int RU = (HG ? 2 : 0) + (HL ? 1 : 0);  // V-value units per halo record
    constexpr int LU = HG ? 2 : 0;                   // unit of L in a record
    extern __shared__ double lds[];
    int64_t i, off = 0;
    int lane = 0, W = 0;
    bool live;       // with a.stage_psum, threads past the last node stay for the block's sums
    int64_t blk;     // tile, or block of kBlock nodes
    if constexpr (TL) {
        const int64_t t = remapped_block();
        if (t >= a.n_tiles) return;  // whole block
        blk = t;
        const int64_t h0 = a.hb_off[t];
        const int nh = (int)(a.hb_off[t + 1] - h0);
        for (int q = threadIdx.x; q < nh; q += kBlock) {
            const int64_t j = a.hb_ids[h0 + q];
            if constexpr (TM == 1) {  // FP32 records
                float* ldsf = reinterpret_cast<float*>(lds);
                double g[V];
                if constexpr (HG) {
                    load_state32<V>(a.G32, 2 * j, g);
                    lds_store_unit32<RU>(ldsf, q, 0, g);
                    load_state32<V>(a.G32, 2 * j + 1, g);
                    lds_store_unit32<RU>(ldsf, q, 1, g);
                }
                if constexpr (HL) {
                    load_state32<V>(a.L32, j, g);
                    lds_store_unit32<RU>(ldsf, q, LU, g);
                }
            } else {
                double g[V];
                if constexpr (HG) {
                    load_state<V>(a.G, 2 * j, g);
                    lds_store_unit<RU>(lds, q, 0, g);
                    load_state<V>(a.G, 2 * j + 1, g);
                    lds_store_unit<RU>(lds, q, 1, g);
                }
                if constexpr (HL) {
                    load_state<V>(a.L, j, g);
                    lds_store_unit<RU>(lds, q, LU, g);
                }
            }
        }
        __syncthreads();
        const int pos = threadIdx.x;
        i = t * kBlock + a.rowb[t * kBlock + pos];
        live = i < a.N;
        if (!live && !a.stage_psum) return;
        if (live) {
            lane = pos & 63;
            const int64_t sl = t * (kBlock / kSellC) + (pos >> 6);
            off = a.sb_off[sl];
            W = a.sb_w[sl];
        }
    } else {
        i = remapped_node();
        if (i - threadIdx.x >= a.N) return;  // whole block
        blk = (i - threadIdx.x) / kBlock;
        live = i < a.N;
        if (!live && !a.stage_psum) return;
        if (live) {
            const int64_t slice = i >> 6;
            lane = (int)(i & 63);
            off = a.sell_off[slice];
            W = a.sell_w[slice];
        }
    }
    double tx[V], ty[V], th[V];
#pragma unroll
    for (int v = 0; v < V; ++v) tx[v] = ty[v] = th[v] = 0.0;
#pragma unroll 1
    for (int e0 = 0; e0 < W; e0 += BT) {
        int32_t r[BT];
        double2 w[BT];
        double wl[BT];
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int e = (e0 + b < W) ? e0 + b : e0;
            const int64_t q = off + (int64_t)e * kSellC + lane;
            if constexpr (TL)
                r[b] = a.lb[q];
            else
                r[b] = a.trow[q];
            if constexpr (HG) {
                if constexpr (TM == 1) {
                    const float2 w32 = TL ? a.twb32[q] : a.tw32[q];
                    w[b] = make_double2((double)w32.x, (double)w32.y);
                } else {
                    w[b] = TL ? a.twb[q] : a.tw[q];
                }
            }
            if constexpr (HL) {
                if constexpr (TM == 1)
                    wl[b] = (double)(TL ? a.twl32[q] : a.tl32[q]);
                else
                    wl[b] = TL ? a.twl[q] : a.tl[q];
            }
        }
        double g0[BT][V], g1[BT][V], lv[BT][V];
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int64_t rb = r[b];
            if constexpr (TM == 1) {
                const float* ldsf = reinterpret_cast<const float*>(lds);
                if constexpr (HG) {
                    if constexpr (TL) {
                        lds_load_unit32<RU>(ldsf, rb, 0, g0[b]);
                        lds_load_unit32<RU>(ldsf, rb, 1, g1[b]);
                    } else {
                        load_state32<V>(a.G32, 2 * rb, g0[b]);
                        load_state32<V>(a.G32, 2 * rb + 1, g1[b]);
                    }
                }
                if constexpr (HL) {
                    if constexpr (TL)
                        lds_load_unit32<RU>(ldsf, rb, LU, lv[b]);
                    else
                        load_state32<V>(a.L32, rb, lv[b]);
                }
            } else {
                if constexpr (HG) {
                    if constexpr (TL) {
                        lds_load_unit<RU>(lds, rb, 0, g0[b]);
                        lds_load_unit<RU>(lds, rb, 1, g1[b]);
                    } else {
                        load_state<V>(a.G, 2 * rb, g0[b]);
                        load_state<V>(a.G, 2 * rb + 1, g1[b]);
                    }
                }
                if constexpr (HL) {
                    if constexpr (TL)
                        lds_load_unit<RU>(lds, rb, LU, lv[b]);
                    else
                        load_state<V>(a.L, rb, lv[b]);
                }
            }
        }
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            if (e0 + b >= W) continue;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if constexpr (TM == 2) {
                    if constexpr (HG) {
                        tx[v] = tx[v] + w[b].x * g0[b][v];
                        ty[v] = ty[v] + w[b].y * g1[b][v];
                    }
                    if constexpr (HL) th[v] = th[v] + wl[b] * lv[b][v];
                } else {
                    if constexpr (HG) {
                        tx[v] = fma(w[b].x, g0[b][v], tx[v]);
                        ty[v] = fma(w[b].y, g1[b][v], ty[v]);
                    }
                    if constexpr (HL) th[v] = fma(wl[b], lv[b][v], th[v]);
                }
            }
        }
    }
    double o[4] = {0.0, 0.0, 0.0, 0.0};  // the stored output (staged stages)
    if (live) {
        double du[V];
        if constexpr (TM == 1)
            load_state32<V>(a.du32, i, du);
        else
            load_state<V>(a.du, i, du);
        if constexpr (TM == 2) {
            // the phase-B sources in keyword order
            for (int sq = a.seq_b; sq != 0; sq >>= 2) {
                const int kind = sq & 3;
                if (kind == kSeqViscosity) {
#pragma unroll
                    for (int v = 0; v < V; ++v) du[v] = (du[v] - tx[v]) - ty[v];
                } else if (kind == kSeqTominec) {
                    const double ng = -a.tom_gamma;
#pragma unroll
                    for (int v = 0; v < V; ++v) du[v] = du[v] + th[v] * ng;
                } else {
                    const double ng = -a.hv_gamma;
                    for (int s = 0; s < a.k; ++s) {
                        const int64_t q = (int64_t)s * a.N + i;
                        double uj[V];
                        load_state<V>(a.u, a.idx_s[q], uj);
                        const double h = a.hv_s[q];
#pragma unroll
                        for (int v = 0; v < V; ++v) du[v] = du[v] + h * (uj[v] * ng);
                    }
                }
            }
        } else {
            if constexpr (HG) {
#pragma unroll
                for (int v = 0; v < V; ++v) du[v] = (du[v] - tx[v]) - ty[v];
            }
            if constexpr (HL) {
#pragma unroll
                for (int v = 0; v < V; ++v) du[v] = fma(-a.tom_gamma, th[v], du[v]);
            }
        }
        mask_du<V>(a, i, du);
        if constexpr (V == 4) {
            if (a.stage_psum)
                emit_staged(a, i, du, o);
            else
                emit<V>(a, i, du);
        } else {
            emit<V>(a, i, du);
        }
    }
    if constexpr (V == 4) {
        if (a.stage_psum) {  // the same for the whole block
            block_sum4(o);
            if (threadIdx.x == 0)
#pragma unroll
                for (int v = 0; v < 4; ++v) a.stage_psum[4 * blk + v] = o[v];
        }
    }
}

// ----------------------------------------- boundary pass on the state
// First boundary pass (rbfsolver.jl:405-408, PointCloudBCs.jl:49-106) over
// entries [e0, e1): Dirichlet u_b <- g(x_b, t), slip u_b <- projection.
// With homog = 1 it is the adjoint of that map: Dirichlet rows zeroed,
// slip momentum projected.
template <int V>
__global__ __launch_bounds__(kBlock) void k_bc_state(const BcArgs b, double* u)
{
    const int e = b.e0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= b.e1) return;
    const int32_t node = b.bidx[e];
    double s[V];
    load_state<V>(u, node, s);
    if (bc_entry<V>(b, e, node, s)) store_state<V>(u, node, s);
}

// ------------------------------------ residual-viscosity reductions
// mean(u) = sum(u) / (V N)  (ode_mean, auxiliary/mpi.jl:40-52) and the
// lexicographic maximum of |u_i - mean| (ode_maximum, mpi.jl:71-82) with
// zeros replaced by eps() (hyperviscosity.jl:305-311).  Each block reduces
// a contiguous node range in a fixed order into one partial; the next
// kernel combines the partials (every block the same way, same bits).
// Sum pass: block b reduces nodes [b*chunk, (b+1)*chunk) in a fixed order.
// With bc.flag set it also applies the first boundary pass to the nodes it
// visits (each node is touched by exactly one thread here).
constexpr int kRed = 4;

__global__ __launch_bounds__(kBlock) void k_rv_sum(int64_t N, int64_t chunk, double* u, double* psum,
                                                   const BcArgs bc)
{
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = min(N, lo + chunk);
    double acc[4] = {0, 0, 0, 0};
    // kRed nodes per thread and trip, all loads issued first
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)kBlock * kRed) {
        double s[kRed][4];
        int32_t ent[kRed];  // boundary entry of the node, -1 if none
#pragma unroll
        for (int r = 0; r < kRed; ++r) {
            const int64_t i = i0 + (int64_t)r * kBlock;
            ent[r] = -1;
            if (i < hi) {
                load_state<4>(u, i, s[r]);
                if (bc.flag) ent[r] = bc.entry[i];
            }
        }
#pragma unroll
        for (int r = 0; r < kRed; ++r) {
            const int64_t i = i0 + (int64_t)r * kBlock;
            if (i >= hi) break;
            if (ent[r] >= 0) {
                if (bc_node<4>(bc, ent[r], (int32_t)i, s[r])) store_state<4>(u, i, s[r]);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[v] += s[r][v];
        }
    }
    block_sum4(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) psum[4 * blockIdx.x + v] = acc[v];
}

// Max pass: mean from the sum pass's partials (recomputed identically in
// every block; block 0 publishes it), then per-block lexicographic maxima
// of |u_i - mean|.
__global__ __launch_bounds__(kBlock) void k_rv_max(int64_t N, int64_t chunk, const double* u,
                                                   const double* psum, int nparts, double* pmax,
                                                   StepScalars* sc)
{
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = min(N, lo + chunk);
    double s[kRed][4];
    auto load_trip = [&](int64_t base) {
#pragma unroll
        for (int r = 0; r < kRed; ++r) {
            const int64_t i = base + (int64_t)r * kBlock;
            if (i < hi) load_state<4>(u, i, s[r]);
        }
    };
    load_trip(lo + threadIdx.x);  // in flight while the mean is reduced
    double mean[4];
    mean_from_partials(psum, nparts, N, mean);
    if (blockIdx.x == 0 && threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) sc->norms[v] = mean[v];
    double best[4] = {-1.0, -1.0, -1.0, -1.0};  // below every |.|
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)kBlock * kRed) {
        if (i0 != lo + threadIdx.x) load_trip(i0);
#pragma unroll
        for (int r = 0; r < kRed; ++r) {
            if (i0 + (int64_t)r * kBlock >= hi) break;
            double c[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) c[v] = fabs(s[r][v] - mean[v]);
            if (lex_less(best, c))
#pragma unroll
                for (int v = 0; v < 4; ++v) best[v] = c[v];
        }
    }
    block_lexmax4(best);
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) pmax[4 * blockIdx.x + v] = best[v];
}

// ------------------------------------------------------ small kernels
__global__ void k_set_time(StepScalars* sc, double t, int success_iter, int set_si)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->t = t;
        if (set_si) sc->success_iter = success_iter;
    }
}

__global__ void k_advance_time(StepScalars* sc, double dt)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->t = sc->t + dt;
        sc->success_iter = sc->success_iter + 1;
    }
}

// HistoryCallback (history.jl:91-152).  Single thread: shift the time
// history, count the points in use and solve the transposed Vandermonde
// system for the derivative weights (partial-pivoting LU, as `A' \ b`).
// One thread.  The step scalars are read once into registers, the
// shifted history and the weights are computed there, and the results are
// stored once: every global access in a dependent chain would cost a
// memory round trip.  With `advance` set it first ends the step as
// k_advance_time does (t += dt, success_iter += 1), saving that launch.
__global__ __launch_bounds__(64) void k_history_weights(StepScalars* sc, int P, int init, int advance, double dt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double th[kMaxHist];
#pragma unroll
    for (int i = 0; i < kMaxHist; ++i) th[i] = init ? 0.0 : sc->time_hist[i];
    const int head = init ? 0 : (sc->hist_head + P - 1) % P;
    double tnow = sc->t;
    int si = sc->success_iter;
    if (advance) {
        tnow = tnow + dt;
        si = si + 1;
        sc->t = tnow;
        sc->success_iter = si;
    }
#pragma unroll
    for (int i = kMaxHist - 1; i >= 1; --i)
        if (i < P) th[i] = th[i - 1];
    th[0] = tnow;
    const int q = (si + 1 < P) ? si + 1 : P;
#pragma unroll
    for (int i = 0; i < kMaxHist; ++i) sc->time_hist[i] = th[i];
    sc->hist_head = head;
    sc->hist_si = si;
    sc->q = (si > 0) ? q : 0;
    if (si <= 0) return;
    // Vandermonde transpose at the scaled times, A(kk, i) = ts_i^kk, and
    // its LU solve with partial pivoting.  Every index is static (loops run
    // to kMaxHist, guarded by q), so the arrays stay in registers.
    constexpr int H = kMaxHist;
    double A[H][H], b[H], ts[H];
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < H; ++i)
        if (i < q) mx = fmax(mx, fabs(th[i]));
    const double scale = 1.0 / mx;
#pragma unroll
    for (int i = 0; i < H; ++i) ts[i] = i < q ? th[i] * scale : 0.0;
    const double te = ts[0];
#pragma unroll
    for (int kk = 0; kk < H; ++kk) {
#pragma unroll
        for (int i = 0; i < H; ++i) {
            // ts^kk, nearly correctly rounded (double-double)
            double hi = 1.0, lo = 0.0;
#pragma unroll
            for (int r = 0; r < kk; ++r) {
                const double ph = hi * ts[i];
                const double pl = fma(hi, ts[i], -ph) + lo * ts[i];
                hi = ph + pl;
                lo = pl - (hi - ph);
            }
            A[kk][i] = hi;
        }
        double hi = 1.0, lo = 0.0;
#pragma unroll
        for (int r = 0; r + 1 < kk; ++r) {
            const double ph = hi * te;
            const double pl = fma(hi, te, -ph) + lo * te;
            hi = ph + pl;
            lo = pl - (hi - ph);
        }
        b[kk] = (kk == 0) ? 0.0 : (double)kk * hi;
    }
    int piv[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        piv[j] = j;
        if (j >= q) continue;
        int p = j;
        double pm = fabs(A[j][j]);
#pragma unroll
        for (int i = j + 1; i < H; ++i)
            if (i < q && fabs(A[i][j]) > pm) {
                pm = fabs(A[i][j]);
                p = i;
            }
        piv[j] = p;
        if (p != j) {
#pragma unroll
            for (int c2 = 0; c2 < H; ++c2) {
                const double ajc = A[j][c2];
                double apc = ajc;
#pragma unroll
                for (int r = j + 1; r < H; ++r)
                    if (r == p) apc = A[r][c2];
                A[j][c2] = apc;
#pragma unroll
                for (int r = j + 1; r < H; ++r)
                    if (r == p) A[r][c2] = ajc;
            }
        }
        const double r = 1.0 / A[j][j];
#pragma unroll
        for (int i = j + 1; i < H; ++i)
            if (i < q) A[i][j] = A[i][j] * r;
#pragma unroll
        for (int c2 = j + 1; c2 < H; ++c2)
#pragma unroll
            for (int i = j + 1; i < H; ++i)
                if (c2 < q && i < q) A[i][c2] = __dsub_rn(A[i][c2], __dmul_rn(A[i][j], A[j][c2]));
    }
#pragma unroll
    for (int j = 0; j < H; ++j)
        if (j < q && piv[j] != j) {
            const double bj = b[j];
            double bp = bj;
#pragma unroll
            for (int r = j + 1; r < H; ++r)
                if (r == piv[j]) bp = b[r];
            b[j] = bp;
#pragma unroll
            for (int r = j + 1; r < H; ++r)
                if (r == piv[j]) b[r] = bj;
        }
#pragma unroll
    for (int j = 0; j < H; ++j)
#pragma unroll
        for (int i = j + 1; i < H; ++i)
            if (j < q && i < q) b[i] = __dsub_rn(b[i], __dmul_rn(A[i][j], b[j]));
#pragma unroll
    for (int j = H - 1; j >= 0; --j) {
        if (j >= q) continue;
        b[j] = b[j] / A[j][j];
#pragma unroll
        for (int i = 0; i < j; ++i) b[i] = __dsub_rn(b[i], __dmul_rn(A[i][j], b[j]));
    }
#pragma unroll
    for (int i = 0; i < H; ++i)
        if (i < q) sc->tw[i] = scale * b[i];
}

// Snapshot u into the ring slot of the most recent entry and rebuild
// approx_du = sum_i w_i u^(i) (history.jl:113-129).
__global__ __launch_bounds__(kBlock) void k_history_apply(int64_t NV, int P, const double* u,
                                                          double* hist, double* approx,
                                                          const StepScalars* sc)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NV) return;
    const int head = sc->hist_head;
    const double val = u[q];
    hist[(int64_t)head * NV + q] = val;
    double acc = 0.0;
    const int nq = sc->q;
    for (int r = 0; r < nq; ++r) {
        const int slot = (head + r) % P;
        const double h = (r == 0) ? val : hist[(int64_t)slot * NV + q];
        acc = fma(sc->tw[r], h, acc);
    }
    approx[q] = acc;
}

// Zhang-Shu limiter (positivity_zhang_shu_point2d.jl:22-83), Jacobi style:
// pass 1 computes the blended state of every node needing it, pass 2
// writes the non-zero ones back.
template <int V>
__device__ __forceinline__ double limiter_variable(int var, double gamma, const double* s)
{
    if (var == MF_VAR_DENSITY || V == 1) return s[0];
    return (gamma - 1.0) * (s[3] - 0.5 * (s[1] * s[1] + s[2] * s[2]) / s[0]);
}

// Stencil entries per load batch of the minimum: the batch's index loads,
// then its neighbour loads, are all in flight together.
constexpr int kLimBatch = 10;

// Node i with stencil minimum vmin below the threshold: the theta-blend with
// the stencil mean (in stencil order) into loc, and the row's flag byte.
template <int V>
__device__ __forceinline__ void limiter_finish(int64_t N, int K, const int32_t* idx, const double* u,
                                               double* loc, uint8_t* flag, double thr, int var, double gamma,
                                               int64_t i, double vmin)
{
    uint8_t f = 0;
    if (vmin < thr) {
        double o[V];
        double um[V];
#pragma unroll
        for (int v = 0; v < V; ++v) um[v] = 0.0;
        for (int s = 0; s < K; ++s) {
            double uj[V];
            load_state<V>(u, idx[(int64_t)s * N + i], uj);
#pragma unroll
            for (int v = 0; v < V; ++v) um[v] = um[v] + uj[v];
        }
#pragma unroll
        for (int v = 0; v < V; ++v) um[v] = um[v] / (double)K;
        const double vmean = limiter_variable<V>(var, gamma, um);
        const double theta = (vmean - thr) / (vmean - vmin);
        double ui[V];
        load_state<V>(u, i, ui);
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = fma(theta, ui[v], (1.0 - theta) * um[v]);
        // only non-zero rows are copied back (a row of zeros stays unwritten)
        bool nz = false;
#pragma unroll
        for (int v = 0; v < V; ++v) nz |= (o[v] != 0.0);
        if (nz) {
            store_state<V>(loc, i, o);
            f = 1;
        }
    }
    flag[i] = f;
}

template <int V>
__global__ __launch_bounds__(kBlock) void k_limiter_blend(int64_t N, int K, const int32_t* idx,
                                                          const double* u, double* loc, uint8_t* flag,
                                                          double thr, int var, double gamma)
{
    const int64_t i = remapped_node();
    if (i >= N) return;
    // the limited variable of every stencil node: density alone needs only
    // the first component of each neighbour
    const bool dens = var == MF_VAR_DENSITY || V == 1;
    double vmin = INFINITY;
#pragma unroll 1
    for (int s0 = 0; s0 < K; s0 += kLimBatch) {
        int32_t j[kLimBatch];
#pragma unroll
        for (int b = 0; b < kLimBatch; ++b) j[b] = (s0 + b < K) ? idx[(int64_t)(s0 + b) * N + i] : (int32_t)i;
        double val[kLimBatch];
        if (dens) {
#pragma unroll
            for (int b = 0; b < kLimBatch; ++b) val[b] = u[(int64_t)j[b] * V];
        } else {
#pragma unroll
            for (int b = 0; b < kLimBatch; ++b) {
                double uj[V];
                load_state<V>(u, j[b], uj);
                val[b] = limiter_variable<V>(var, gamma, uj);
            }
        }
#pragma unroll
        for (int b = 0; b < kLimBatch; ++b)
            if (s0 + b < K) vmin = (val[b] < vmin) ? val[b] : vmin;
    }
    limiter_finish<V>(N, K, idx, u, loc, flag, thr, var, gamma, i, vmin);
}

// The same with the tile's stencil halo in LDS (the tiles of the stencil
// kernel): the limited variable of every halo node is evaluated once into
// LDS and the minimum reads it through the 16-bit halo positions la.  Same
// values, same order, same bits.
template <int V>
__global__ __launch_bounds__(kBlock) void k_limiter_blend_tl(int64_t N, int K, const int32_t* idx,
                                                             const uint16_t* la, int64_t n_tiles,
                                                             const int64_t* ha_off, const int32_t* ha_ids,
                                                             const double* u, double* loc, uint8_t* flag,
                                                             double thr, int var, double gamma)
{
    extern __shared__ double lvar[];
    const int64_t t = remapped_block();
    if (t >= n_tiles) return;  // whole block
    const bool dens = var == MF_VAR_DENSITY || V == 1;
    const int64_t h0 = ha_off[t];
    const int nh = (int)(ha_off[t + 1] - h0);
    for (int q = threadIdx.x; q < nh; q += kBlock) {
        const int32_t j = ha_ids[h0 + q];
        if (dens) {
            lvar[q] = u[(int64_t)j * V];
        } else {
            double uj[V];
            load_state<V>(u, j, uj);
            lvar[q] = limiter_variable<V>(var, gamma, uj);
        }
    }
    __syncthreads();
    const int64_t i = t * kBlock + threadIdx.x;
    if (i >= N) return;
    double vmin = INFINITY;
#pragma unroll 1
    for (int s0 = 0; s0 < K; s0 += kLimBatch) {
        double val[kLimBatch];
#pragma unroll
        for (int b = 0; b < kLimBatch; ++b) val[b] = lvar[(s0 + b < K) ? la[tm_pos(s0 + b, i, K)] : 0];
#pragma unroll
        for (int b = 0; b < kLimBatch; ++b)
            if (s0 + b < K) vmin = (val[b] < vmin) ? val[b] : vmin;
    }
    limiter_finish<V>(N, K, idx, u, loc, flag, thr, var, gamma, i, vmin);
}

template <int V>
__global__ __launch_bounds__(kBlock) void k_limiter_commit(int64_t N, const double* loc, const uint8_t* flag,
                                                           double* u)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || !flag[i]) return;
    double o[V];
    load_state<V>(loc, i, o);
    store_state<V>(u, i, o);
}

__global__ void k_axpby(int64_t n, double a, const double* x, double b, double* y)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] = fma(a, x[q], b * y[q]);
}

__global__ void k_fill(double* p, int64_t n, double v)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

// ------------------------------------------------------------ IGR source
// SourceIGR (sources/IGR.jl): sigma from 20 conjugate-gradient iterations on
//   (rho^-1 - alpha (Dx rho^-1 Dx + Dy rho^-1 Dy)) sigma = b,
//   b = alpha (trace^2 + trace_sq) from the primitive-variable derivatives,
// then du_m1 -= Dx sigma, du_m2 -= Dy sigma.  Scalar node fields; every
// reduction is per block (fixed node ranges) into partials, which the next
// kernel's blocks each sum in the same fixed order, so every block, and
// every run, gets the same bits.  Three launches per CG iteration over the
// LDS tiles, four with global gathers.
struct IgrArgs {
    int64_t N;
    int k;
    int64_t nblk;
    const int32_t* __restrict__ idx;
    const double2* __restrict__ wxy;
    double alpha;
    int maxiter;
    const double* __restrict__ u;
    double *rinv, *b, *x, *r, *p, *c, *t1, *t2;
    double* p2;       // the tile path's second search direction (ping-pong with p)
    double* part_rr;  // per-block partial sums of r . r
    double* part_pc;  // and of p . c
    IgrScalars* sc;
    // the stencil kernel's LDS tiles (tile t = nodes t*kBlock ..; halo ids
    // and 16-bit halo positions of every stencil entry)
    int64_t n_tiles;
    const int64_t* __restrict__ ha_off;
    const int32_t* __restrict__ ha_ids;
    const uint16_t* __restrict__ la;
};

// Block-wide sum of one value per thread (every thread gets it).
__device__ __forceinline__ double block_sum1(double v)
{
    double w[4] = {v, 0.0, 0.0, 0.0};
    block_sum4(w);
    return w[0];
}

// Sum of the per-block partials in a fixed order (every block the same bits).
__device__ __forceinline__ double sum_partials(const double* part, int64_t n)
{
    double acc = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += kBlock) acc += part[q];
    return block_sum1(acc);
}

// Sums over the stencil of node i of Dx times the field fx and Dy times the
// field fy (each stencil entry's weights read once).
template <int BT>
__device__ __forceinline__ double2 igr_grad2(const IgrArgs& a, int64_t i, const double* fx, const double* fy)
{
    double sx = 0.0, sy = 0.0;
#pragma unroll 1
    for (int s0 = 0; s0 < a.k; s0 += BT) {
        int32_t j[BT];
        double2 w[BT];
#pragma unroll
        for (int q = 0; q < BT; ++q) {
            const int s = s0 + q < a.k ? s0 + q : s0;
            j[q] = a.idx[(int64_t)s * a.N + i];
            w[q] = a.wxy[(int64_t)s * a.N + i];
        }
        double gx[BT], gy[BT];
#pragma unroll
        for (int q = 0; q < BT; ++q) {
            gx[q] = fx[j[q]];
            gy[q] = fx == fy ? gx[q] : fy[j[q]];
        }
#pragma unroll
        for (int q = 0; q < BT; ++q) {
            if (s0 + q >= a.k) continue;
            sx = fma(w[q].x, gx[q], sx);
            sy = fma(w[q].y, gy[q], sy);
        }
    }
    return make_double2(sx, sy);
}

// (Dx f, Dy f) at node i.
template <int BT>
__device__ __forceinline__ double2 igr_grad(const IgrArgs& a, int64_t i, const double* __restrict__ f)
{
    return igr_grad2<BT>(a, i, f, f);
}

// The same sums with the tile's halo values of the gathered field(s) in
// LDS (one double, or (fx, fy) as one 16-byte record, per halo entry),
// addressed by the 16-bit halo positions: same values in the same order.
template <int BT, bool TWO>
__device__ __forceinline__ double2 igr_grad_tl(const IgrArgs& a, int64_t i, const double* lds)
{
    double sx = 0.0, sy = 0.0;
#pragma unroll 1
    for (int s0 = 0; s0 < a.k; s0 += BT) {
        int pos[BT];
        double2 w[BT];
#pragma unroll
        for (int q = 0; q < BT; ++q) {
            const int s = s0 + q < a.k ? s0 + q : s0;
            const int64_t tq = tm_pos(s, i, a.k);  // tile-major (la, and wxy the tile-major copy)
            pos[q] = a.la[tq];
            w[q] = a.wxy[tq];
        }
        double gx[BT], gy[BT];
#pragma unroll
        for (int q = 0; q < BT; ++q) {
            if constexpr (TWO) {
                const double2 v = reinterpret_cast<const double2*>(lds)[pos[q]];
                gx[q] = v.x;
                gy[q] = v.y;
            } else {
                gx[q] = gy[q] = lds[pos[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < BT; ++q) {
            if (s0 + q >= a.k) continue;
            sx = fma(w[q].x, gx[q], sx);
            sy = fma(w[q].y, gy[q], sy);
        }
    }
    return make_double2(sx, sy);
}

// Stages the halo of tile t: f0 (and f1 beside it with TWO) of every halo node.
template <bool TWO>
__device__ __forceinline__ void igr_stage(const IgrArgs& a, int64_t t, const double* f0, const double* f1,
                                          double* lds)
{
    const int64_t h0 = a.ha_off[t];
    const int nh = (int)(a.ha_off[t + 1] - h0);
    for (int q = threadIdx.x; q < nh; q += kBlock) {
        const int32_t j = a.ha_ids[h0 + q];
        if constexpr (TWO)
            reinterpret_cast<double2*>(lds)[q] = make_double2(f0[j], f1[j]);
        else
            lds[q] = f0[j];
    }
    __syncthreads();
}

// b, 1/rho, sigma = 0, r = b, p = 0 (update_igr_rhs!, IGR.jl:117-157, and
// update_sigma!, :163-180), and the partial sums of r^2.  The trace is
// D_x rho + D_y v1 as the reference writes it (flux_x[1] + flux_y[2]).
__global__ __launch_bounds__(kBlock) void k_igr_init(const IgrArgs a)
{
    const int64_t lb = remapped_block();
    if (lb >= a.nblk) return;  // whole block
    const int64_t i = lb * kBlock + threadIdx.x;
    double rr = 0.0;
    if (i < a.N) {
        double gxr = 0.0, gx1 = 0.0, gx2 = 0.0, gy1 = 0.0, gy2 = 0.0;
        for (int s = 0; s < a.k; ++s) {
            const int64_t q = (int64_t)s * a.N + i;
            const double2 w = a.wxy[q];
            double uj[4];
            load_state<4>(a.u, a.idx[q], uj);
            const double v1 = uj[1] / uj[0], v2 = uj[2] / uj[0];
            gxr = fma(w.x, uj[0], gxr);
            gx1 = fma(w.x, v1, gx1);
            gx2 = fma(w.x, v2, gx2);
            gy1 = fma(w.y, v1, gy1);
            gy2 = fma(w.y, v2, gy2);
        }
        const double trace = gxr + gy1;
        const double tsq = gx1 * gx1 + 2.0 * gy1 * gx2 + gy2 * gy2;
        const double bi = a.alpha * (trace * trace + tsq);
        double ui[4];
        load_state<4>(a.u, i, ui);
        a.rinv[i] = 1.0 / ui[0];
        a.b[i] = bi;
        a.x[i] = 0.0;
        a.r[i] = bi;
        a.p[i] = 0.0;
        rr = bi * bi;
    }
    rr = block_sum1(rr);
    if (threadIdx.x == 0) a.part_rr[lb] = rr;
}

// Start of CG iteration `it` (IterativeSolvers' cg! iterate): the residual
// norm res = |r| from the partials, the tolerance sqrt(eps) |r0| at it = 0,
// stop once res <= tol (the host launches at most maxiter iterations);
// else p = r + beta p with beta = res^2 / prev^2 (prev = 1 at it = 0).
// Block 0 publishes res (slot it & 1, where the next iteration finds it as
// prev), the tolerance and the stop flag for this iteration's kernels.
__global__ __launch_bounds__(kBlock) void k_igr_p(const IgrArgs a, int it)
{
    const double res = sqrt(sum_partials(a.part_rr, a.nblk));
    const double tol = it == 0 ? sqrt(DBL_EPSILON) * res : a.sc->tol;
    const double prev = it == 0 ? 1.0 : a.sc->res[(it - 1) & 1];
    const bool done = res <= tol;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.sc->res[it & 1] = res;
        a.sc->tol = tol;
        a.sc->done = done ? 1 : 0;
    }
    if (done) return;
    const int64_t i = remapped_node();
    if (i >= a.N) return;
    const double beta = res * res / (prev * prev);
    a.p[i] = a.r[i] + beta * a.p[i];
}

// t = rho^-1 (Dx p, Dy p).
template <int BT>
__global__ __launch_bounds__(kBlock) void k_igr_t(const IgrArgs a)
{
    if (a.sc->done) return;
    const int64_t i = remapped_node();
    if (i >= a.N) return;
    const double2 g = igr_grad<BT>(a, i, a.p);
    const double ri = a.rinv[i];
    a.t1[i] = ri * g.x;
    a.t2[i] = ri * g.y;
}

// k_igr_p and k_igr_t in one launch over the LDS tiles: every block forms
// res, beta and the stop test from the partials as k_igr_p does (block 0
// publishes them), stages r + beta p_old of its halo nodes (the same
// expression as the new direction of those nodes), writes its own nodes'
// new direction to p_new (another buffer: neighbouring blocks still read
// p_old) and forms t.
template <int BT>
__global__ __launch_bounds__(kBlock) void k_igr_pt_tl(const IgrArgs a, const double* __restrict__ p_old,
                                                      double* __restrict__ p_new, int it)
{
    extern __shared__ double lds[];
    const double res = sqrt(sum_partials(a.part_rr, a.nblk));
    const double tol = it == 0 ? sqrt(DBL_EPSILON) * res : a.sc->tol;
    const double prev = it == 0 ? 1.0 : a.sc->res[(it - 1) & 1];
    const bool done = res <= tol;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.sc->res[it & 1] = res;
        a.sc->tol = tol;
        a.sc->done = done ? 1 : 0;
    }
    if (done) return;
    const int64_t t = remapped_block();
    if (t >= a.n_tiles) return;  // whole block
    const double beta = res * res / (prev * prev);
    const int64_t h0 = a.ha_off[t];
    const int nh = (int)(a.ha_off[t + 1] - h0);
    for (int q = threadIdx.x; q < nh; q += kBlock) {
        const int32_t j = a.ha_ids[h0 + q];
        lds[q] = a.r[j] + beta * p_old[j];
    }
    __syncthreads();
    const int64_t i = t * kBlock + threadIdx.x;
    if (i >= a.N) return;
    p_new[i] = a.r[i] + beta * p_old[i];
    const double2 g = igr_grad_tl<BT, false>(a, i, lds);
    const double ri = a.rinv[i];
    a.t1[i] = ri * g.x;
    a.t2[i] = ri * g.y;
}

// c = rho^-1 p - alpha (Dx t1 + Dy t2), and partial sums of p . c.
template <int BT>
__global__ __launch_bounds__(kBlock) void k_igr_c(const IgrArgs a)
{
    if (a.sc->done) return;
    const int64_t lb = remapped_block();
    if (lb >= a.nblk) return;  // whole block
    const int64_t i = lb * kBlock + threadIdx.x;
    double pc = 0.0;
    if (i < a.N) {
        const double2 g = igr_grad2<BT>(a, i, a.t1, a.t2);  // (Dx t1, Dy t2)
        const double ci = a.rinv[i] * a.p[i] - a.alpha * (g.x + g.y);
        a.c[i] = ci;
        pc = a.p[i] * ci;
    }
    pc = block_sum1(pc);
    if (threadIdx.x == 0) a.part_pc[lb] = pc;
}

// k_igr_c over the LDS tiles (tile t is block t of k_igr_c).
template <int BT>
__global__ __launch_bounds__(kBlock) void k_igr_c_tl(const IgrArgs a)
{
    extern __shared__ double lds[];
    if (a.sc->done) return;
    const int64_t lb = remapped_block();
    if (lb >= a.nblk) return;  // whole block
    igr_stage<true>(a, lb, a.t1, a.t2, lds);
    const int64_t i = lb * kBlock + threadIdx.x;
    double pc = 0.0;
    if (i < a.N) {
        const double2 g = igr_grad_tl<BT, true>(a, i, lds);  // (Dx t1, Dy t2)
        const double ci = a.rinv[i] * a.p[i] - a.alpha * (g.x + g.y);
        a.c[i] = ci;
        pc = a.p[i] * ci;
    }
    pc = block_sum1(pc);
    if (threadIdx.x == 0) a.part_pc[lb] = pc;
}

// alpha = res^2 / (p . c) from the partials, sigma += alpha p,
// r -= alpha c, and partial sums of r^2.
__global__ __launch_bounds__(kBlock) void k_igr_x(const IgrArgs a, int it)
{
    if (a.sc->done) return;
    const int64_t lb = remapped_block();
    if (lb >= a.nblk) return;  // whole block
    const double pc = sum_partials(a.part_pc, a.nblk);
    const double res = a.sc->res[it & 1];
    const int64_t i = lb * kBlock + threadIdx.x;
    double rr = 0.0;
    if (i < a.N) {
        const double al = res * res / pc;
        a.x[i] += al * a.p[i];
        const double ri = a.r[i] - al * a.c[i];
        a.r[i] = ri;
        rr = ri * ri;
    }
    rr = block_sum1(rr);
    if (threadIdx.x == 0) a.part_rr[lb] = rr;
}

// du_m1 -= Dx sigma, du_m2 -= Dy sigma (the flux_igr terms, IGR.jl:225-237),
// then the second boundary pass and the output of the rhs! or stage.
template <int BT>
__global__ __launch_bounds__(kBlock) void k_igr_apply(const IgrArgs g, const RhsArgs a)
{
    const int64_t i = remapped_node();
    if (i >= g.N) return;
    const double2 s = igr_grad<BT>(g, i, g.x);
    double du[4];
    load_state<4>(a.du, i, du);
    du[1] -= s.x;
    du[2] -= s.y;
    mask_du<4>(a, i, du);
    emit<4>(a, i, du);
}

static IgrArgs make_igr(mf_ctx* c, const double* u)
{
    IgrArgs g{};
    g.N = c->N;
    g.k = c->k;
    g.nblk = (c->N + kBlock - 1) / kBlock;
    g.idx = c->d_idx;
    g.wxy = c->d_wxy;
    g.alpha = c->igr_alpha;
    g.maxiter = c->igr_maxiter;
    g.u = u;
    double* f = c->d_igr;
    const int64_t N = c->N;
    g.rinv = f;
    g.b = f + N;
    g.x = f + 2 * N;
    g.r = f + 3 * N;
    g.p = f + 4 * N;
    g.c = f + 5 * N;
    g.t1 = f + 6 * N;
    g.t2 = f + 7 * N;
    g.p2 = f + 8 * N;
    g.part_rr = c->d_igr_part;
    g.part_pc = c->d_igr_part + g.nblk;
    g.n_tiles = c->n_tiles;
    g.ha_off = c->d_ha_off;
    g.ha_ids = c->d_ha_ids;
    g.la = c->d_la;
    g.sc = c->d_igr_sc;
    return g;
}

// ============================================================ launchers
static RhsArgs make_args(mf_ctx* c, double* u, double* du)
{
    RhsArgs a{};
    a.N = c->N;
    a.k = c->k;
    a.idx = c->d_idx;
    a.didx = c->d_didx;
    a.wxy = c->d_wxy;
    a.hv = c->d_hv;
    a.wxy32 = c->d_wxy32;
    a.hv32 = c->d_hv32;
    a.hv_gamma = c->hv_gamma;
    a.eq.a1 = c->p0;
    a.eq.a2 = c->p1;
    a.eq.gamma = c->p0;
    if (c->eq == MF_EQ_EULER_2D) a.eq.a1 = a.eq.a2 = 0.0;
    a.u = u;
    a.du = du;
    a.flag = c->have_bc ? c->bc.flag : nullptr;
    a.entry = c->have_bc ? c->bc.entry : nullptr;
    a.bnormal = c->have_bc ? c->bc.bnormal : nullptr;
    a.visc = c->visc_mode;
    a.hv_first = c->hv_first;
    a.c_rv = c->c_rv;
    a.c_uw = c->c_uw;
    a.h = c->h_visc;
    a.approx = c->d_approx;
    a.sc = c->d_sc;
    a.idx_s = c->d_idx_s;
    a.wx_s = c->d_wx_s;
    a.wy_s = c->d_wy_s;
    a.hv_s = c->d_hv_s;
    a.pmax = c->d_partials ? c->d_partials + 4 * c->n_psum : nullptr;
    a.nparts = c->n_pmax;
    a.norms_out = c->d_sc ? c->d_sc->norms : nullptr;
    a.G = c->d_G;
    a.G32 = reinterpret_cast<float*>(c->d_G);
    a.du32 = c->d_du32;
    a.eps = c->d_eps;
    a.epsc = c->d_epsc;
    a.sell_off = c->d_sell_off;
    a.sell_w = c->d_sell_w;
    a.trow = c->d_trow;
    a.tw = c->d_tw;
    a.tw32 = c->d_tw32;
    a.n_tiles = c->n_tiles;
    a.ha_off = c->d_ha_off;
    a.ha_ids = c->d_ha_ids;
    a.la = c->d_la;
    a.la_s = c->d_la_s;
    a.hb_off = c->d_hb_off;
    a.hb_ids = c->d_hb_ids;
    a.sb_off = c->d_sb_off;
    a.sb_w = c->d_sb_w;
    a.rowb = c->d_rowb;
    a.lb = c->d_lb;
    a.twb = c->d_twb;
    a.twb32 = c->d_twb32;
    a.lap = c->d_lap;
    a.lap32 = c->d_lap32;
    a.lap_s = c->d_lap_s;
    a.tom_gamma = c->tom_gamma;
    a.L = c->d_L;
    a.L32 = reinterpret_cast<float*>(c->d_L);
    a.tl = c->d_tl;
    a.tl32 = c->d_tl32;
    a.twl = c->d_twl;
    a.twl32 = c->d_twl32;
    // the active sources in keyword order
    const bool fly = c->have_hv && c->hv_gamma != 0.0;
    const bool tom = c->have_lap && c->tom_gamma != 0.0;
    const bool rv = c->visc_mode != MF_VISC_NONE;
    int order[3], n = 0;
    for (int k : {MF_SRC_FLYER, MF_SRC_TOMINEC, MF_SRC_VISCOSITY})
        if ((k == MF_SRC_FLYER && fly) || (k == MF_SRC_TOMINEC && tom) || (k == MF_SRC_VISCOSITY && rv))
            order[n++] = k;
    std::sort(order, order + n, [&](int x, int y) { return c->src_rank[x] < c->src_rank[y]; });
    a.flyer_a = fly && order[0] == MF_SRC_FLYER;
    a.seq_b = 0;
    int sh = 0;
    for (int q = 0; q < n; ++q) {
        if (order[q] == MF_SRC_FLYER && a.flyer_a) continue;
        const int code = order[q] == MF_SRC_VISCOSITY ? kSeqViscosity
                                                      : (order[q] == MF_SRC_TOMINEC ? kSeqTominec : kSeqFlyer);
        a.seq_b |= code << sh;
        sh += 2;
    }
    return a;
}

static BcArgs make_bc(mf_ctx* c, double dt_off, int homog)
{
    BcArgs b{};
    b.bidx = c->bc.bidx;
    b.btag = c->bc.btag;
    b.bnormal = c->bc.bnormal;
    b.btable = c->bc.btable;
    b.px = c->d_x;
    b.py = c->d_y;
    b.gamma = c->p0;
    b.sc = c->d_sc;
    b.flag = c->bc.flag;
    b.entry = c->bc.entry;
    b.dt_off = dt_off;
    b.homog = homog;
    b.tag_kind = c->bc.d_kind;
    b.tag_fn = c->bc.d_fn;
    b.tag_params = c->bc.d_params;
    b.ecode = c->bc.ecode;
    return b;
}

// One boundary pass on a state buffer; tags run one after another when a
// node belongs to several (the reference applies them in tag order).
// Launch on the context stream.  While the context is being timed
// (c->probe set), the dispatch of probe slot `slot` carries its own start
// and stop events (hipExtLaunchKernelGGL), which time the kernel alone.
// Slots: 0 stencil kernel, 1 transposed gather, 2 sum pass, 3 max pass,
// 4 first boundary pass, 5 second boundary pass.
template <typename... KArgs, typename... Args>
static void launch_k(mf_ctx* c, int slot, void (*kernel)(KArgs...), dim3 G, dim3 B, size_t lds, Args... args)
{
    if (c->probe && slot >= 0) {
        c->probe_used |= 1u << slot;
        hipExtLaunchKernelGGL(kernel, G, B, (uint32_t)lds, c->stream, c->probe[2 * slot], c->probe[2 * slot + 1], 0,
                              args...);
    } else {
        hipLaunchKernelGGL(kernel, G, B, lds, c->stream, args...);
    }
}

static int bc_state_pass(mf_ctx* c, double* u, double dt_off, int homog, bool slip_only)
{
    if (!c->have_bc || c->bc.nb == 0) return MF_OK;
    if (slip_only && !c->bc.any_slip) return MF_OK;
    BcArgs b = make_bc(c, dt_off, homog);
    if (slip_only) b.tag_kind = c->bc.d_kind_slip;
    auto launch = [&](int e0, int e1) {
        if (e1 <= e0) return;
        b.e0 = e0;
        b.e1 = e1;
        const unsigned g = (unsigned)((e1 - e0 + kBlock - 1) / kBlock);
        const int slot = homog ? -1 : (slip_only ? 5 : 4);
        if (c->nvar == 4)
            launch_k(c, slot, k_bc_state<4>, dim3(g), dim3(kBlock), 0, b, u);
        else
            launch_k(c, slot, k_bc_state<1>, dim3(g), dim3(kBlock), 0, b, u);
    };
    if (c->bc.multi) {
        for (int g = 0; g < c->bc.ntags; ++g) launch(c->bc.tag_ptr[g], c->bc.tag_ptr[g + 1]);
    } else {
        launch(0, c->bc.nb);
    }
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// Stencil batch of the Tominec variants (compiled for one batch size only).
constexpr int def_bt(int ns) { return ns ? 5 : 4; }

// hvk bit 0: Flyer source, bit 1: Tominec source.
#define MF_LAUNCH_HVK(KERNEL, ...)                                                      \
    do {                                                                                \
        if (hvk & 2) {                                                                  \
            if constexpr (V == 4 && BT == def_bt(NS)) {                                 \
                if (hvk == 3 && rv) launch_k(c, 0, KERNEL<V, NS, 3, true, __VA_ARGS__>, G, B, lds, a);  \
                else if (hvk == 3) launch_k(c, 0, KERNEL<V, NS, 3, false, __VA_ARGS__>, G, B, lds, a);  \
                else if (rv) launch_k(c, 0, KERNEL<V, NS, 2, true, __VA_ARGS__>, G, B, lds, a);         \
                else launch_k(c, 0, KERNEL<V, NS, 2, false, __VA_ARGS__>, G, B, lds, a);                \
            }                                                                           \
        } else if (hvk && rv) launch_k(c, 0, KERNEL<V, NS, 1, true, __VA_ARGS__>, G, B, lds, a);         \
        else if (hvk) launch_k(c, 0, KERNEL<V, NS, 1, false, __VA_ARGS__>, G, B, lds, a);                \
        else if (rv) launch_k(c, 0, KERNEL<V, NS, 0, true, __VA_ARGS__>, G, B, lds, a);                  \
        else launch_k(c, 0, KERNEL<V, NS, 0, false, __VA_ARGS__>, G, B, lds, a);                         \
    } while (0)

template <int V, int NS, int BT, bool TL>
static void launch_f64(mf_ctx* c, const RhsArgs& a, int hvk, bool rv, const dim3& G, size_t lds)
{
    const dim3 B(kBlock);
    MF_LAUNCH_HVK(k_rhs_f64, BT, TL);
}

template <int V, int NS, int BT, bool D16, bool TL>
static void launch_s32(mf_ctx* c, const RhsArgs& a, int hvk, bool rv, const dim3& G, size_t lds)
{
    const dim3 B(kBlock);
    MF_LAUNCH_HVK(k_rhs_s32, D16, BT, TL);
}

template <int V, int NS, int BT, bool TL>
static void launch_ref(mf_ctx* c, const RhsArgs& a, int hvk, bool rv, const dim3& G, size_t lds)
{
    const dim3 B(kBlock);
    MF_LAUNCH_HVK(k_rhs_ref, BT, TL);
}
#undef MF_LAUNCH_HVK

// The operators in the tile-major layout of the tile kernels.
static RhsArgs tile_major_args(const mf_ctx* c, const RhsArgs& a)
{
    RhsArgs t = a;
    t.wxy = c->d_wxy_t;
    t.wxy32 = c->d_wxy32_t;
    t.hv = c->d_hv_t;
    t.hv32 = c->d_hv32_t;
    t.lap = c->d_lap_t;
    t.lap32 = c->d_lap32_t;
    t.wx_s = c->d_wx_s_t;
    t.wy_s = c->d_wy_s_t;
    t.hv_s = c->d_hv_s_t;
    t.lap_s = c->d_lap_s_t;
    return t;
}

template <int V, int NS, int BT>
static void launch_flux_bt(mf_ctx* c, int prec, const RhsArgs& a_in, int hvk, bool rv, unsigned grid)
{
    const dim3 G(grid);
    // the tile kernels read the tile-major operator copies
    const bool tl = c->tiles_a && !c->tm_dirty;
    const RhsArgs a = tl ? tile_major_args(c, a_in) : a_in;
    const size_t lds = tl ? (size_t)c->ha_max * rec_len<V>() * sizeof(double) : 0;  // halo records
    if (prec == MF_PREC_F64) {
        if (tl)
            launch_f64<V, NS, BT, true>(c, a, hvk, rv, G, lds);
        else
            launch_f64<V, NS, BT, false>(c, a, hvk, rv, G, 0);
    } else if (prec == MF_PREC_F64_REF) {
        if (tl)
            launch_ref<V, NS, BT, true>(c, a, hvk, rv, G, lds);
        else
            launch_ref<V, NS, BT, false>(c, a, hvk, rv, G, 0);
    } else if (tl) {
        launch_s32<V, NS, BT, false, true>(c, a, hvk, rv, G, lds);
    } else if (c->didx_ok) {
        launch_s32<V, NS, BT, true, false>(c, a, hvk, rv, G, 0);
    } else {
        launch_s32<V, NS, BT, false, false>(c, a, hvk, rv, G, 0);
    }
}

// Stencil batch sizes compiled per stencil size: the tuning value picks
// one (it must divide the stencil size; 0 = the default, the first).  The
// Tominec variants exist for the default only.
template <int V>
static void launch_flux_v(mf_ctx* c, int prec, const RhsArgs& a, int hvk, bool rv, unsigned grid)
{
    const int bt = (hvk & 2) ? 0 : c->tune_flux_batch[prec];
    switch (c->k) {
        case 15:
            if (bt == 3) launch_flux_bt<V, 15, 3>(c, prec, a, hvk, rv, grid);
            else launch_flux_bt<V, 15, 5>(c, prec, a, hvk, rv, grid);
            break;
        case 20:
            if (bt == 10) launch_flux_bt<V, 20, 10>(c, prec, a, hvk, rv, grid);
            else launch_flux_bt<V, 20, 5>(c, prec, a, hvk, rv, grid);
            break;
        case 30:
            if (bt == 3) launch_flux_bt<V, 30, 3>(c, prec, a, hvk, rv, grid);
            else if (bt == 6) launch_flux_bt<V, 30, 6>(c, prec, a, hvk, rv, grid);
            else if (bt == 10) launch_flux_bt<V, 30, 10>(c, prec, a, hvk, rv, grid);
            else launch_flux_bt<V, 30, 5>(c, prec, a, hvk, rv, grid);
            break;
        default: launch_flux_bt<V, 0, 4>(c, prec, a, hvk, rv, grid); break;
    }
}

// gt bit 0: G of the viscosity source, bit 1: L of the Tominec source (its
// variants exist for the default batch only).
template <int TM, bool TL>
static void launch_transpose_tl(mf_ctx* c, const RhsArgs& a, unsigned grid, size_t lds, int gt)
{
    const dim3 G(grid), B(kBlock);
    if (gt == 2) {
        launch_k(c, 1, k_rhs_transpose<4, TM, 4, TL, 2>, G, B, lds, a);
    } else if (gt == 3) {
        launch_k(c, 1, k_rhs_transpose<4, TM, 4, TL, 3>, G, B, lds, a);
    } else {
        switch (c->tune_transpose_batch[TM]) {
            case 2: launch_k(c, 1, k_rhs_transpose<4, TM, 2, TL, 1>, G, B, lds, a); break;
            case 8: launch_k(c, 1, k_rhs_transpose<4, TM, 8, TL, 1>, G, B, lds, a); break;
            default: launch_k(c, 1, k_rhs_transpose<4, TM, 4, TL, 1>, G, B, lds, a); break;
        }
    }
}

template <int TM>
static void launch_transpose(mf_ctx* c, const RhsArgs& a, unsigned grid, int gt)
{
    const int ru = ((gt & 1) ? 2 : 0) + ((gt & 2) ? 1 : 0);  // 4-value units per halo record
    if (c->tiles_b && c->tiles_b_rec == 4 * ru)
        launch_transpose_tl<TM, true>(c, a, grid,
                                      (size_t)c->hb_max * 4 * ru * (TM == 1 ? sizeof(float) : sizeof(double)), gt);
    else
        launch_transpose_tl<TM, false>(c, a, grid, 0, gt);
}

// The IGR source on state u: CG for sigma, then the flux terms added to du
// (written without output by the kernels before) and the output a.
static int run_igr(mf_ctx* c, const double* u, const RhsArgs& a, unsigned grid)
{
    const IgrArgs g = make_igr(c, u);
    const dim3 G(grid), B(kBlock);
    launch_k(c, -1, k_igr_init, G, B, 0, g);
    // the stencil kernel's tiles, when built (one tile per block of nodes)
    const bool tl = c->tiles_a && c->d_la && c->n_tiles == g.nblk && !c->tm_dirty && c->d_wxy_t;
    const size_t lds1 = sizeof(double) * (size_t)c->ha_max, lds2 = 2 * lds1;
    IgrArgs gt = g;  // the tile kernels read the tile-major copy of (Dx, Dy)
    gt.wxy = c->d_wxy_t;
    for (int it = 0; it < c->igr_maxiter; ++it) {
        if (tl) {
            // three launches per iteration: the direction update rides on
            // the t pass, p and p2 alternate (p is zero after k_igr_init)
            double* p_old = (it & 1) ? g.p2 : g.p;
            double* p_new = (it & 1) ? g.p : g.p2;
            IgrArgs gi = gt;
            gi.p = p_new;
            IgrArgs gx = g;
            gx.p = p_new;
            launch_k(c, -1, k_igr_pt_tl<5>, G, B, lds1, gt, (const double*)p_old, p_new, it);
            launch_k(c, -1, k_igr_c_tl<5>, G, B, lds2, gi);
            launch_k(c, -1, k_igr_x, G, B, 0, gx, it);
            continue;
        }
        launch_k(c, -1, k_igr_p, G, B, 0, g, it);
        launch_k(c, -1, k_igr_t<5>, G, B, 0, g);
        launch_k(c, -1, k_igr_c<5>, G, B, 0, g);
        launch_k(c, -1, k_igr_x, G, B, 0, g, it);
    }
    launch_k(c, -1, k_igr_apply<5>, G, B, 0, g, a);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

int run_rhs(mf_ctx* c, int prec, double* d_u, double* d_du, double dt_off, const StageOut& st)
{
    return run_rhs_parts(c, prec, d_u, d_du, dt_off, st, RHS_ALL);
}

int run_rhs_parts(mf_ctx* c, int prec, double* d_u, double* d_du, double dt_off, const StageOut& st,
                  int parts)
{
    if (!c->have_flux) return fail(c, MF_ERR_STATE, "rhs!: no flux operator (mf_compute_operator/mf_set_operator)");
    if (prec != MF_PREC_F64 && prec != MF_PREC_SPLIT32 && prec != MF_PREC_F64_REF)
        return fail(c, MF_ERR_ARGUMENT, "rhs!: unknown precision");
    const bool hv = c->have_hv && c->hv_gamma != 0.0;
    const bool tom = c->have_lap && c->tom_gamma != 0.0;
    const bool rv = c->visc_mode != MF_VISC_NONE;
    if (rv && c->nvar != 4) return fail(c, MF_ERR_UNSUPPORTED, "viscosity sources need the Euler equations");
    // the residual viscosity's norms and the IGR source's conjugate-gradient
    // dot products are sums over the whole point set: on a grouped set they
    // would couple the independent clouds
    if (c->n_groups > 1 && (c->visc_mode == MF_VISC_RESIDUAL || c->igr_alpha != 0.0))
        return fail(c, MF_ERR_UNSUPPORTED,
                    "rhs!: the residual viscosity and IGR sources reduce over all points; not on grouped point sets");
    if (tom && c->nvar != 4) return fail(c, MF_ERR_UNSUPPORTED, "the Tominec source needs the Euler equations");
    if ((rv || tom) && !c->have_T) return fail(c, MF_ERR_STATE, "viscosity source without transposed operator");
    if (tom && !c->lap_T) return fail(c, MF_ERR_STATE, "Tominec source: transposed lap not prepared");
    if (tom && rv && c->visc_mode == MF_VISC_RESIDUAL &&
        c->src_rank[MF_SRC_TOMINEC] < c->src_rank[MF_SRC_VISCOSITY])
        return fail(c, MF_ERR_UNSUPPORTED,
                    "a Tominec source before the residual viscosity source (its du would need lap'(lap u) first)");
    const bool igr = c->igr_alpha != 0.0;
    if (igr && c->nvar != 4) return fail(c, MF_ERR_UNSUPPORTED, "the IGR source needs the Euler equations");
    if (igr && !c->d_igr) return fail(c, MF_ERR_STATE, "IGR source: buffers not prepared");
    if (igr && ((hv && c->src_rank[MF_SRC_IGR] < c->src_rank[MF_SRC_FLYER]) ||
                (tom && c->src_rank[MF_SRC_IGR] < c->src_rank[MF_SRC_TOMINEC]) ||
                (rv && c->src_rank[MF_SRC_IGR] < c->src_rank[MF_SRC_VISCOSITY])))
        return fail(c, MF_ERR_UNSUPPORTED, "the IGR source must come after the other sources");
    const int hvk = (hv ? 1 : 0) | (tom ? 2 : 0);
    // first boundary pass on u
    // with the residual viscosity the first boundary pass rides on the sum pass
    const bool fused_bc = rv && c->visc_mode == MF_VISC_RESIDUAL && c->have_bc && !c->bc.multi;
    // a stage handed over by the previous rhs! (stepping) has had this
    // rhs!'s first boundary pass and carries its sums
    const int64_t nblk = (c->N + kBlock - 1) / kBlock;
    if (st.sums_in || st.sums_out) {
        if (c->visc_mode != MF_VISC_RESIDUAL || igr || !c->d_stage_psum || (c->have_bc && c->bc.multi) ||
            (c->tiles_b && c->n_tiles != nblk))
            return fail(c, MF_ERR_STATE, "rhs!: stage sums without the residual viscosity's transposed kernel");
    }
    int rc = ((parts & RHS_BC) && !fused_bc && !st.sums_in) ? bc_state_pass(c, d_u, dt_off, 0, false) : MF_OK;
    if (rc) return rc;
    // with the IGR source the other kernels leave du (masked) for its last
    // kernel, which writes the output
    RhsArgs out_args = make_args(c, d_u, d_du);
    out_args.st_kind = st.kind;
    out_args.dt = st.dt;
    out_args.uprev = st.uprev;
    out_args.out = st.out;
    if (st.sums_out) {
        out_args.stage_psum = c->d_stage_psum;
        out_args.obc = make_bc(c, st.out_dt_off, 0);
        out_args.obc.t_adv = st.out_t_adv;
        if (!c->have_bc) out_args.obc.flag = nullptr;
    }
    RhsArgs a = out_args;
    if (igr) a.st_kind = ST_NONE;
    const unsigned grid = xcd_grid(nblk);
    if (rv || tom) {
        if (rv && c->visc_mode == MF_VISC_RESIDUAL && (parts & RHS_REDUCE)) {
            const int ns = c->n_psum, nm = c->n_pmax;
            if (st.sums_in) {
                launch_k(c, 3, k_rv_max, dim3(nm), dim3(kBlock), 0, c->N, (c->N + nm - 1) / nm, (const double*)d_u,
                         (const double*)c->d_stage_psum, (int)nblk, c->d_partials + 4 * ns, c->d_sc);
            } else {
                BcArgs bc = make_bc(c, dt_off, 0);
                if (!((parts & RHS_BC) && fused_bc)) bc.flag = nullptr;
                launch_k(c, 2, k_rv_sum, dim3(ns), dim3(kBlock), 0, c->N, (c->N + ns - 1) / ns, d_u,
                         c->d_partials, bc);
                launch_k(c, 3, k_rv_max, dim3(nm), dim3(kBlock), 0, c->N, (c->N + nm - 1) / nm, (const double*)d_u,
                         (const double*)c->d_partials, ns, c->d_partials + 4 * ns, c->d_sc);
            }
        }
        RhsArgs pa = a;
        pa.st_kind = ST_NONE;
        pa.stage_psum = nullptr;
        if (parts & RHS_FLUX) launch_flux_v<4>(c, prec, pa, hvk, rv, grid);
        if (parts & RHS_TRANSPOSE) {
            const int gt = (rv ? 1 : 0) | (tom ? 2 : 0);
            if (prec == MF_PREC_F64)
                launch_transpose<0>(c, a, grid, gt);
            else if (prec == MF_PREC_SPLIT32)
                launch_transpose<1>(c, a, grid, gt);
            else
                launch_transpose<2>(c, a, grid, gt);
        }
    } else {
        if (parts & RHS_FLUX) {
            if (c->nvar == 4)
                launch_flux_v<4>(c, prec, a, hvk, false, grid);
            else
                launch_flux_v<1>(c, prec, a, hvk, false, grid);
        }
    }
    MF_HIP(c, hipGetLastError());
    if (igr && (parts & RHS_IGR) && (rc = run_igr(c, d_u, out_args, grid))) return rc;
    // second boundary pass on u: only the slip projection changes anything
    rc = (parts & RHS_BC) ? bc_state_pass(c, d_u, dt_off, 0, true) : MF_OK;
    return rc;
}

int run_set_time(mf_ctx* c, double t, int success_iter)
{
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, c->stream, c->d_sc, t, success_iter,
                       success_iter >= 0 ? 1 : 0);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

int run_advance_time(mf_ctx* c, double dt)
{
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, c->stream, c->d_sc, dt);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

int run_history(mf_ctx* c, const double* d_u, int init, double advance_dt)
{
    const int P = c->approx_order + 1;
    hipLaunchKernelGGL(k_history_weights, dim3(1), dim3(1), 0, c->stream, c->d_sc, P, init,
                       advance_dt != 0.0 ? 1 : 0, advance_dt);
    const int64_t NV = c->N * c->nvar;
    hipLaunchKernelGGL(k_history_apply, dim3((unsigned)((NV + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, NV, P, d_u, c->d_hist, c->d_approx, c->d_sc);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

int run_limiter(mf_ctx* c, double* d_u, double thr, int var)
{
    const int64_t nblk = (c->N + kBlock - 1) / kBlock;
    const unsigned grid = xcd_grid(nblk);
    const double gam = c->p0;
    // blended rows, then one flag byte per node (the buffer's tail)
    uint8_t* flag = reinterpret_cast<uint8_t*>(c->d_lim + (int64_t)c->nvar * c->N);
    if (c->tiles_a && c->d_la) {
        // the stencil kernel's tiles: one LDS double per halo entry
        const size_t lds = sizeof(double) * (size_t)c->ha_max;
        const unsigned tg = xcd_grid(c->n_tiles);
        if (c->nvar == 4)
            hipLaunchKernelGGL(k_limiter_blend_tl<4>, dim3(tg), dim3(kBlock), lds, c->stream, c->N, c->k, c->d_idx,
                               c->d_la, c->n_tiles, c->d_ha_off, c->d_ha_ids, d_u, c->d_lim, flag, thr, var, gam);
        else
            hipLaunchKernelGGL(k_limiter_blend_tl<1>, dim3(tg), dim3(kBlock), lds, c->stream, c->N, c->k, c->d_idx,
                               c->d_la, c->n_tiles, c->d_ha_off, c->d_ha_ids, d_u, c->d_lim, flag, thr, var, gam);
        if (c->nvar == 4)
            hipLaunchKernelGGL(k_limiter_commit<4>, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream, c->N,
                               c->d_lim, flag, d_u);
        else
            hipLaunchKernelGGL(k_limiter_commit<1>, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream, c->N,
                               c->d_lim, flag, d_u);
    } else if (c->nvar == 4) {
        hipLaunchKernelGGL(k_limiter_blend<4>, dim3(grid), dim3(kBlock), 0, c->stream, c->N, c->k,
                           c->d_idx, d_u, c->d_lim, flag, thr, var, gam);
        hipLaunchKernelGGL(k_limiter_commit<4>, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream,
                           c->N, c->d_lim, flag, d_u);
    } else {
        hipLaunchKernelGGL(k_limiter_blend<1>, dim3(grid), dim3(kBlock), 0, c->stream, c->N, c->k,
                           c->d_idx, d_u, c->d_lim, flag, thr, var, gam);
        hipLaunchKernelGGL(k_limiter_commit<1>, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream,
                           c->N, c->d_lim, flag, d_u);
    }
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

int run_axpby(mf_ctx* c, int64_t n, double a, const double* x, double b, double* y)
{
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, n, a, x, b, y);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

int run_fill(mf_ctx* c, double* p, int64_t n, double v)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, p, n, v);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// exported for the adjoint source
int bc_pass_homog(mf_ctx* c, double* u) { return bc_state_pass(c, u, 0.0, 1, false); }
int bc_pass_state(mf_ctx* c, double* u, double dt_off) { return bc_state_pass(c, u, dt_off, 0, false); }

}  // namespace mf
