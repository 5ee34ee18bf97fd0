// Batched RBF-FD weights by the null-space (projector) form of the saddle
// system, with the dense products on FP64 matrix cores.
//
// Reference: compute_flux_operator, pointcloudsolver/compute_operators.jl:
// 409-453 (first derivatives) and :549-594 (k-th derivatives): per stencil,
// shift_stencil (:225-246), the blocks R = phi(|X_i - X_j|) and P (:191-223),
// the right-hand sides L phi at the mirrored points and L p at the centre
// (:248-263, :283-291, :345-384), M \ rhs with M = [R P; P' 0] (:435), the
// scaling by s_d^k (:438-439, :579-580).  The reference factors M by
// Bunch-Kaufman; this kernel solves the same system another way:
//
//   P = Q U            (Cholesky QR: G = P'P = U'U, Q = P U^-1, n x np)
//   w = w_p + y,       w_p = Q U'^-1 c  (so that P'w = c)
//   K y = f,           K = s Pi R Pi + Q Q'  (SPD),  f = s Pi (b - R w_p)
//
// with Pi = I - QQ' and s = (-1)^ceil(p/2), the sign that makes the
// polyharmonic spline r^p conditionally positive definite on the null space
// of P'.  K is formed as s R - W Q' - Q W' with X = R Q, Y = Q'X and
// W = s (X - Q Y/2) - Q/2, and solved by Gauss-Jordan elimination without
// pivoting (K is symmetric positive definite).  The products G, X, Y, W,
// Q'h and the rank-2np update of K run as v_mfma_f64_16x16x4 tiles.
//
// One wavefront (workgroup) per stencil, np <= 16 monomials, n > np.  The
// kernel is built for up to MT 16-row tiles: MT = 2 (n <= 32: 128 VGPRs and
// under 10 KiB of LDS, four wavefronts per SIMD), MT = 3 (n <= 48: 160
// VGPRs, three) and MT = 4 (n <= 64: 192 VGPRs, two); a ragged batch runs
// as one launch per variant over the stencils of its size class.  (A persistent grid that loops over
// stencils took twice the registers: the loop-invariant LDS addresses of
// the unrolled phases stay live across the loop.)
#include <hip/hip_runtime.h>

#include "mf_device.h"

namespace mf {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kNsLz = 18;  // row stride (doubles) of the 16 x 16 LDS tile

// Row stride of the n x np LDS matrices: odd, so that the 16 rows one MFMA
// operand reads fall in distinct banks.
constexpr int ns_lq(int np) { return np <= 11 ? 11 : 17; }

struct NsArgs {
    const double* x;
    const double* y;
    const int64_t* ptr;  // ragged stencils (or null: ELL)
    const int32_t* ridx;
    const int32_t* sid;  // stencil of each workgroup (null: workgroup b runs stencil b)
    const int32_t* ell;  // ELL, slot-major (ptr null)
    int ell_k;
    int64_t N;
    int p, np, kd;
    int rows;            // LDS rows of the n x np matrices (>= largest n, >= 16)
    int lq;              // their row stride, ns_lq(np)
    double kfact, sgn;
    PhsTable tab;
    double* wx;
    double* wy;
    int* status;
};

// D = A B + C for 16 x 4 by 4 x 16 FP64 tiles.  Lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and holds D[4 r + l / 16][l % 16] in element r.
__device__ __forceinline__ d4 mfma4(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double bcast(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// 1/sqrt(d) and 1/d for d > 0: the hardware estimates (about 2^-29
// relative) refined by two Newton steps each, about 1 ulp; cheaper than
// the correctly rounded quotient and root, whose last bit is not needed here
__device__ __forceinline__ double rsqrt_nr(double d)
{
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

__device__ __forceinline__ double rcp_nr(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    y = y * fma(-d, y, 2.0);
    y = y * fma(-d, y, 2.0);
    return y;
}

__device__ __forceinline__ double wave_absmax(double v)
{
    v = fabs(v);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// phi(|X_i - X_j|) = r^p from the scaled coordinates in LDS (interleaved x,
// y).  r = sqrt(d2) as d2 y with y the Newton-refined 1/sqrt(d2), then one
// Newton step on r itself (r += y/2 (d2 - r^2), residual by fma): within an
// ulp, without the scaling and special-case steps of the library root (the
// coordinates are scaled to [-1, 1], so 0 < d2 <= 8 off the diagonal);
// r = 0 on the diagonal.  P > 0: the power at compile time (the products of
// ipow, without its loop); P = 0: p at run time.
template <int P>
__device__ __forceinline__ double phs_r(const double* co, int i, int j, int p)
{
#pragma clang fp contract(off)
    const double dx = co[2 * i] - co[2 * j], dy = co[2 * i + 1] - co[2 * j + 1];
    const double d2 = dx * dx + dy * dy;
    double r = 0.0;
    if (d2 > 0.0) {
        const double y = rsqrt_nr(d2);
        const double r0 = d2 * y;
        r = fma(0.5 * y, fma(-r0, r0, d2), r0);
    }
    if constexpr (P == 0) {
        return ipow(r, p);
    } else {
        double v = r;
#pragma unroll
        for (int e = 1; e < P; ++e) v = v * r;
        return v;
    }
}

// NF, NPF: the stencil size and monomial count when the launch is compiled
// for one shape (every stencil of it has NF nodes), 0 when they are taken
// at run time; a compiled shape folds the size guards and unrolls every
// loop over nodes and monomials.
template <int MT, int P, int NF, int NPF>
__global__ __launch_bounds__(64, MT == 2 ? 4 : (MT == 3 ? 3 : 1)) void k_weights_ns(const NsArgs a)
{
#pragma clang fp contract(off)
    constexpr int NL = 16 * MT;  // largest stencil
    extern __shared__ double sm[];
    const int lane = threadIdx.x;
    const int q = lane >> 4, c = lane & 15;
    const int LQ = NPF ? ns_lq(NPF) : a.lq, LZ = kNsLz;
    const int rows = NF ? (NF < 16 ? 16 : NF) : a.rows;
    double* co = sm;                      // NL x (x, y)
    double* Qm = sm + 2 * NL;             // rows x LQ: P, then Q (columns < np)
    double* Xm = Qm + rows * LQ;          // rows x LQ: X = R Q, then W (columns < np)
    double* Zm = Xm + rows * LQ;          // 16 x LZ: G, L of G, Z, then K tiles
    double* col = Zm + 16 * LZ;           // NL: broadcast column (16-byte aligned)
    double* hv = col + NL;                // 2 NL: h = b - X g (2 per node)
    double* gs = hv + 2 * NL;             // 32: g (2 per monomial)
    double* ss = gs + 32;                 // 32: s = Q'h (2 per monomial)
    double* dg = ss + 32;                 // 16: 1 / L_jj of the Cholesky factor of G
    const int np = NPF ? NPF : a.np;

    {
        const int64_t e = a.sid ? (int64_t)a.sid[blockIdx.x] : (int64_t)blockIdx.x;
        int64_t off;
        int n;
        if (a.ptr) {
            off = a.ptr[e];
            n = NF ? NF : (int)(a.ptr[e + 1] - off);
        } else {
            off = e;
            n = NF ? NF : a.ell_k;
        }
        auto sidx = [&](int j) -> int32_t { return a.ptr ? a.ridx[off + j] : a.ell[(int64_t)j * a.N + e]; };
        const int T = (n + 15) >> 4;      // 16-row tiles
        const int nk = (n + 3) >> 2;      // k-steps over the n nodes
        const int npk = (np + 3) >> 2;    // k-steps over the monomials
        bool ok = true;

        // ---- shift_stencil (:225-246): centre on node 0, scale per dimension
        const int32_t c0 = sidx(0);
        double px = 0.0, py = 0.0;
        if (lane < n) {
            const int32_t qn = sidx(lane);
            px = a.x[qn] - a.x[c0];
            py = a.y[qn] - a.y[c0];
        }
        const double sx = 1.0 / wave_absmax(px), sy = 1.0 / wave_absmax(py);
        const double xs = px * sx, ys = py * sy;
        if (lane < NL) {
            co[2 * lane] = xs;
            co[2 * lane + 1] = ys;
        }

        // ---- P row of this node (poly_block, :211-223), into LDS
        if (lane < rows) {
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                double v = 0.0;
                if (t < np && lane < n) {
                    int ax, ay;
                    mono_exp(t, &ax, &ay);
                    v = ipow(xs, ax) * ipow(ys, ay);
                }
                if (t < LQ) Qm[lane * LQ + t] = v;
            }
        }
        __syncthreads();

        // ---- G = P'P on the matrix cores
        {
            d4 g = {0.0, 0.0, 0.0, 0.0};
            for (int ks = 0; ks < nk; ++ks) {
                const int k = 4 * ks + q;
                const double v = (k < n && c < np) ? Qm[k * LQ + c] : 0.0;
                g = mfma4(v, v, g);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) Zm[(4 * r + q) * LZ + c] = g[r];
        }
        __syncthreads();

        // ---- G = L L' (lane = row of G), L kept in Zm
        {
            double gr[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) gr[t] = lane < 16 ? Zm[lane * LZ + t] : 0.0;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (j < np) {
                    const double d = bcast(gr[j], j);
                    ok = ok && d > 0.0;
                    const double rs = rsqrt_nr(d);
                    if (lane == 0) dg[j] = rs;
                    const double l = lane >= j ? gr[j] * rs : 0.0;
                    if (lane < NL) col[lane] = l;
                    __syncthreads();
#pragma unroll
                    for (int cc = j + 1; cc < 16; ++cc) gr[cc] = fma(-l, col[cc], gr[cc]);
                    gr[j] = l;
                    __syncthreads();
                }
            }
            if (lane < 16) {
#pragma unroll
                for (int t = 0; t < 16; ++t) Zm[lane * LZ + t] = t <= lane && lane < np ? gr[t] : 0.0;
            }
            __syncthreads();
        }

        // ---- Q = P L'^-1, row by row in place
        if (lane < n) {
            double qr[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                qr[t] = 0.0;
                if (t < np) {
                    double aq = Qm[lane * LQ + t];
#pragma unroll
                    for (int s = 0; s < t; ++s) aq = fma(-qr[s], Zm[t * LZ + s], aq);
                    qr[t] = aq * dg[t];
                }
            }
#pragma unroll
            for (int t = 0; t < 16; ++t)
                if (t < np) Qm[lane * LQ + t] = qr[t];
        }
        // g = L^-1 c, lane = monomial (c: L applied to the monomials at the
        // centre, poly_linearoperator :283-291)
        {
            int ax, ay;
            mono_exp(lane & 15, &ax, &ay);
            double a0 = (lane < np && ax == a.kd && ay == 0) ? a.kfact : 0.0;
            double a1 = (lane < np && ay == a.kd && ax == 0) ? a.kfact : 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                if (t < np) {
                    const double rd = dg[t];
                    const double g0 = bcast(a0, t) * rd, g1 = bcast(a1, t) * rd;
                    if (lane == t) {
                        a0 = g0;
                        a1 = g1;
                    } else if (lane > t && lane < 16) {
                        const double l = Zm[lane * LZ + t];
                        a0 = fma(-l, g0, a0);
                        a1 = fma(-l, g1, a1);
                    }
                }
            }
            if (lane < 16) {
                gs[2 * lane] = lane < np ? a0 : 0.0;
                gs[2 * lane + 1] = lane < np ? a1 : 0.0;
            }
        }
        __syncthreads();

        // ---- X = R Q on the matrix cores.  For n <= 32 the A-operand values
        //      R[16 rt + c][4 ks + q] are kept: by the symmetry of R they are
        //      the C-layout values R[16 it + 4 r + q][16 jt + c] the tiles of K
        //      start from (rt = jt, ks = 4 it + r), which saves evaluating R twice.
        double ra[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) ra[t] = 0.0;
        if (MT == 2 || T <= 2) {
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
                if (rt < T) {
                    const int i = 16 * rt + c;
                    d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int ks = 0; ks < 8; ++ks) {
                        if (ks < nk) {
                            const int k = 4 * ks + q;
                            const double av = (i < n && k < n) ? phs_r<P>(co, i, k, a.p) : 0.0;
                            const double bv = (k < n && c < np) ? Qm[k * LQ + c] : 0.0;
                            ra[8 * rt + ks] = av;
                            acc = mfma4(av, bv, acc);
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * rt + 4 * r + q;
                        if (row < n && c < np) Xm[row * LQ + c] = acc[r];
                    }
                }
            }
        } else {
#pragma unroll
            for (int rt = 0; rt < MT; ++rt) {
                if (rt < T) {
                    const int i = 16 * rt + c;
                    d4 acc = {0.0, 0.0, 0.0, 0.0};
                    for (int ks = 0; ks < nk; ++ks) {
                        const int k = 4 * ks + q;
                        const double av = (i < n && k < n) ? phs_r<P>(co, i, k, a.p) : 0.0;
                        const double bv = (k < n && c < np) ? Qm[k * LQ + c] : 0.0;
                        acc = mfma4(av, bv, acc);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * rt + 4 * r + q;
                        if (row < n && c < np) Xm[row * LQ + c] = acc[r];
                    }
                }
            }
        }
        __syncthreads();

        // ---- right-hand sides (mirror_stencil :248-263, rbf_linearoperator
        //      :345-358): b at the mirrored points, centre at (eps, eps);
        //      w_p = Q g; h = b - X g
        double wp0 = 0.0, wp1 = 0.0, h0 = 0.0, h1 = 0.0;
        if (lane < n) {
            double mx = -xs, my = -ys;
            if (lane == 0) mx = my = DBL_EPSILON;
            const double b0 = phs_eval(a.tab, mx, mx, my);
            const double b1 = phs_eval(a.tab, my, mx, my);
            double x0 = 0.0, x1 = 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                if (t < np) {
                    const double gt0 = gs[2 * t], gt1 = gs[2 * t + 1];
                    const double xv = Xm[lane * LQ + t], qv = Qm[lane * LQ + t];
                    wp0 = fma(qv, gt0, wp0);
                    wp1 = fma(qv, gt1, wp1);
                    x0 = fma(xv, gt0, x0);
                    x1 = fma(xv, gt1, x1);
                }
            }
            h0 = b0 - x0;
            h1 = b1 - x1;
        }
        if (lane < NL) {
            hv[2 * lane] = h0;
            hv[2 * lane + 1] = h1;
        }
        __syncthreads();

        // ---- s = Q'h (matrix cores; columns 0 and 1 of the tile)
        {
            d4 acc = {0.0, 0.0, 0.0, 0.0};
            for (int ks = 0; ks < nk; ++ks) {
                const int k = 4 * ks + q;
                const double av = (k < n && c < np) ? Qm[k * LQ + c] : 0.0;
                const double bv = (k < n && c < 2) ? hv[2 * k + c] : 0.0;
                acc = mfma4(av, bv, acc);
            }
            if (c < 2) {
#pragma unroll
                for (int r = 0; r < 4; ++r) ss[2 * (4 * r + q) + c] = acc[r];
            }
        }
        __syncthreads();

        // f = sgn (h - Q s), lane = node
        double f0 = 0.0, f1 = 0.0;
        if (lane < n) {
            double a0 = h0, a1 = h1;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                if (t < np) {
                    const double qv = Qm[lane * LQ + t];
                    a0 = fma(-qv, ss[2 * t], a0);
                    a1 = fma(-qv, ss[2 * t + 1], a1);
                }
            }
            f0 = a.sgn * a0;
            f1 = a.sgn * a1;
        }

        // ---- Y = Q'X, Z = sgn Y / 2 + I / 2 (matrix cores)
        {
            d4 acc = {0.0, 0.0, 0.0, 0.0};
            for (int ks = 0; ks < nk; ++ks) {
                const int k = 4 * ks + q;
                const double av = (k < n && c < np) ? Qm[k * LQ + c] : 0.0;
                const double bv = (k < n && c < np) ? Xm[k * LQ + c] : 0.0;
                acc = mfma4(av, bv, acc);
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; ++r)
                Zm[(4 * r + q) * LZ + c] = 0.5 * a.sgn * acc[r] + (4 * r + q == c ? 0.5 : 0.0);
        }
        __syncthreads();

        // ---- W = sgn X - Q Z over the X rows (matrix cores)
#pragma unroll
        for (int rt = 0; rt < MT; ++rt) {
            if (rt < T) {
                const int i = 16 * rt + c;
                d4 acc;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * rt + 4 * r + q;
                    acc[r] = (row < n && c < np) ? a.sgn * Xm[row * LQ + c] : 0.0;
                }
                for (int ks = 0; ks < npk; ++ks) {
                    const int k = 4 * ks + q;
                    const double av = (i < n && k < np) ? -Qm[i * LQ + k] : 0.0;
                    const double bv = Zm[k * LZ + c];
                    acc = mfma4(av, bv, acc);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * rt + 4 * r + q;
                    if (row < n && c < np) Xm[row * LQ + c] = acc[r];
                }
            }
        }
        __syncthreads();

        // ---- K = sgn R - W Q' - Q W' tile by tile (matrix cores), each tile
        //      staged through Zm into the rows of the lanes that own them
        double kr[NL];
#pragma unroll
        for (int t = 0; t < NL; ++t) kr[t] = 0.0;
#pragma unroll
        for (int it = 0; it < MT; ++it) {
            if (it < T) {
#pragma unroll
                for (int jt = 0; jt < MT; ++jt) {
                    if (jt < T) {
                        d4 acc;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = 16 * it + 4 * r + q, cl = 16 * jt + c;
                            double rv;
                            if (it < 2 && jt < 2)
                                rv = (MT == 2 || T <= 2) ? ra[(8 * jt + 4 * it + r) & 15] : phs_r<P>(co, row, cl, a.p);
                            else
                                rv = phs_r<P>(co, row, cl, a.p);
                            acc[r] = (row < n && cl < n) ? a.sgn * rv : (row == cl ? 1.0 : 0.0);
                        }
                        const int ia = 16 * it + c, jb = 16 * jt + c;
                        for (int ks = 0; ks < npk; ++ks) {  // -W Q'
                            const int k = 4 * ks + q;
                            const double av = (ia < n && k < np) ? -Xm[ia * LQ + k] : 0.0;
                            const double bv = (jb < n && k < np) ? Qm[jb * LQ + k] : 0.0;
                            acc = mfma4(av, bv, acc);
                        }
                        for (int ks = 0; ks < npk; ++ks) {  // -Q W'
                            const int k = 4 * ks + q;
                            const double av = (ia < n && k < np) ? -Qm[ia * LQ + k] : 0.0;
                            const double bv = (jb < n && k < np) ? Xm[jb * LQ + k] : 0.0;
                            acc = mfma4(av, bv, acc);
                        }
                        __syncthreads();
#pragma unroll
                        for (int r = 0; r < 4; ++r) Zm[(4 * r + q) * LZ + c] = acc[r];
                        __syncthreads();
                        if ((lane >> 4) == it) {
#pragma unroll
                            for (int t = 0; t < 16; ++t) kr[16 * jt + t] = Zm[c * LZ + t];
                        }
                    }
                }
            }
        }
        __syncthreads();

        // ---- K y = f by Gauss-Jordan elimination (lane = row).  Step j
        //      eliminates column j from every other row; the trailing block
        //      stays symmetric, so row j's entries right of j are column j's.
        double piv = 1.0;  // 1 / pivot of this lane's row
#pragma unroll
        for (int jc = 0; jc < MT; ++jc) {
            if (jc < T) {
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) {
                    const int j = 16 * jc + jj;
                    if (j < n) {
                        if (lane < NL) col[lane] = kr[j];
                        __syncthreads();
                        const double pj = bcast(kr[j], j);
                        const double fj0 = bcast(f0, j), fj1 = bcast(f1, j);
                        ok = ok && pj > 0.0;
                        const double ip = rcp_nr(pj);
                        const double m = lane == j ? 0.0 : kr[j] * ip;
                        if (lane == j) piv = ip;
#pragma unroll
                        for (int cc = jc; cc < MT; ++cc) {
                            if (cc < T) {
#pragma unroll
                                for (int ci = 0; ci < 16; ci += 2) {
                                    const int cl = 16 * cc + ci;
                                    if (cl + 1 > j) {
                                        const double2 cv = *reinterpret_cast<const double2*>(col + cl);
                                        if (cl > j) kr[cl] = fma(-m, cv.x, kr[cl]);
                                        kr[cl + 1] = fma(-m, cv.y, kr[cl + 1]);
                                    }
                                }
                            }
                        }
                        f0 = fma(-m, fj0, f0);
                        f1 = fma(-m, fj1, f1);
                        __syncthreads();
                    }
                }
            }
        }

        // ---- w = w_p + y, scaled back (:438-439, :579-580)
        if (lane < n) {
            const double w0 = wp0 + f0 * piv, w1 = wp1 + f1 * piv;
            const double fx = pow_k(sx, a.kd), fy = pow_k(sy, a.kd);
            const int64_t o = a.ptr ? off + lane : (int64_t)lane * a.N + e;
            a.wx[o] = fx * w0;
            a.wy[o] = fy * w1;
        }
        if (!ok && lane == 0) atomicOr(a.status, 2);
        __syncthreads();
    }
}

// LDS doubles of one wavefront of the MT variant for stencils of up to
// `rows` nodes.
static size_t ns_lds_doubles(int mt, int rows, int lq)
{
    const size_t nl = 16 * (size_t)mt;
    return 2 * nl + 2 * (size_t)rows * lq + 16 * kNsLz + nl + 2 * nl + 32 + 32 + 16;
}

int launch_weights_ns(mf_ctx* c, const int64_t* d_ptr, const int32_t* d_idx, int ell_k, const int32_t* d_sid,
                      int64_t n_launch, int maxn, int uniform_n, int p, int deg, int kd, double* d_wx, double* d_wy,
                      int* d_status)
{
    if (n_launch == 0) return MF_OK;
    NsArgs a;
    a.x = c->d_x;
    a.y = c->d_y;
    a.ptr = d_ptr;
    a.ridx = d_ptr ? d_idx : nullptr;
    a.sid = d_sid;
    a.ell = d_ptr ? nullptr : d_idx;
    a.ell_k = ell_k;
    a.N = c->N;
    a.p = p;
    a.np = (deg + 1) * (deg + 2) / 2;
    a.kd = kd;
    if (a.np > 16) return fail(c, MF_ERR_UNSUPPORTED, "null-space weights: polynomial degree above 4");
    if (maxn > 64) return fail(c, MF_ERR_UNSUPPORTED, "null-space weights: stencil above 64 nodes");
    if ((p & 1) == 0) return fail(c, MF_ERR_UNSUPPORTED, "null-space weights: even PHS power");
    if (n_launch > 0x7fffffff) return fail(c, MF_ERR_ARGUMENT, "null-space weights: too many stencils");
    a.rows = maxn < 16 ? 16 : maxn;
    double kf = 1.0;
    for (int i = 2; i <= kd; ++i) kf *= i;
    a.kfact = kf;
    a.sgn = (((p + 1) / 2) & 1) ? -1.0 : 1.0;  // (-1)^ceil(p/2)
    a.tab = phs_table(p, kd);
    a.wx = d_wx;
    a.wy = d_wy;
    a.status = d_status;
    a.lq = ns_lq(a.np);
    const int mt = maxn <= 32 ? 2 : (maxn <= 48 ? 3 : 4);
    const size_t lds = sizeof(double) * ns_lds_doubles(mt, a.rows, a.lq);
    // the configurations' PHS powers at compile time, others at run time;
    // two stencil shapes compiled in full: 15 nodes with the 6 monomials of
    // degree 2, and 30 nodes with the 10 of degree 3
    void (*kern)(NsArgs) = nullptr;
    auto pick = [&](auto k2, auto k3, auto k4) { kern = mt == 2 ? k2 : (mt == 3 ? k3 : k4); };
    const int shape = uniform_n == 15 && a.np == 6 ? 1 : (uniform_n == 30 && a.np == 10 ? 2 : 0);
    // the larger variants with the 10 monomials of degree 3 compiled in (the
    // stencil size at run time)
    const bool np10 = a.np == 10 && mt > 2;
    switch (p) {
        case 3:
            pick(k_weights_ns<2, 3, 0, 0>, k_weights_ns<3, 3, 0, 0>, k_weights_ns<4, 3, 0, 0>);
            if (np10) kern = mt == 3 ? k_weights_ns<3, 3, 0, 10> : k_weights_ns<4, 3, 0, 10>;
            if (shape) kern = shape == 1 ? k_weights_ns<2, 3, 15, 6> : k_weights_ns<2, 3, 30, 10>;
            break;
        case 5:
            pick(k_weights_ns<2, 5, 0, 0>, k_weights_ns<3, 5, 0, 0>, k_weights_ns<4, 5, 0, 0>);
            if (np10) kern = mt == 3 ? k_weights_ns<3, 5, 0, 10> : k_weights_ns<4, 5, 0, 10>;
            if (shape) kern = shape == 1 ? k_weights_ns<2, 5, 15, 6> : k_weights_ns<2, 5, 30, 10>;
            break;
        case 7:
            pick(k_weights_ns<2, 7, 0, 0>, k_weights_ns<3, 7, 0, 0>, k_weights_ns<4, 7, 0, 0>);
            if (np10) kern = mt == 3 ? k_weights_ns<3, 7, 0, 10> : k_weights_ns<4, 7, 0, 10>;
            if (shape) kern = shape == 1 ? k_weights_ns<2, 7, 15, 6> : k_weights_ns<2, 7, 30, 10>;
            break;
        default: pick(k_weights_ns<2, 0, 0, 0>, k_weights_ns<3, 0, 0, 0>, k_weights_ns<4, 0, 0, 0>); break;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)n_launch), dim3(64), lds, c->stream, a);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

}  // namespace mf
