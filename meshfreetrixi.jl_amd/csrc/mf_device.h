// Device-side helpers shared by the HIP sources: node-major state access,
// the physical fluxes, boundary value functions and launch index mapping.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "mf_internal.h"

namespace mf {

constexpr int kBlock = 256;

// Entry s of node i's stencil in the tile-major layout of the LDS-tile
// kernels: the K slots of each tile of kBlock nodes back to back, kBlock
// entries each.  A tile's entries of one array are then one contiguous run
// (a few pages) instead of K runs N entries apart, which on large point
// sets made the address translation of the slot-major layout the limit.
__device__ __forceinline__ int64_t tm_pos(int s, int64_t i, int K)
{
    return ((i / kBlock) * K + s) * kBlock + (i % kBlock);
}

// Logical node of this thread under the XCD-aware block remap (grid size a
// multiple of 8; workgroup b runs on XCD b % 8).
__device__ __forceinline__ int64_t remapped_node()
{
    const unsigned b = blockIdx.x, per = gridDim.x >> 3;
    const unsigned lb = (b & 7u) * per + (b >> 3);
    return (int64_t)lb * blockDim.x + threadIdx.x;
}

template <int V>
__device__ __forceinline__ void load_state(const double* __restrict__ u, int64_t j, double* s)
{
    if constexpr (V == 4) {
        const double4 q = reinterpret_cast<const double4*>(u)[j];
        s[0] = q.x;
        s[1] = q.y;
        s[2] = q.z;
        s[3] = q.w;
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) s[v] = u[j * V + v];
    }
}

template <int V>
__device__ __forceinline__ void store_state(double* __restrict__ u, int64_t j, const double* s)
{
    if constexpr (V == 4) {
        reinterpret_cast<double4*>(u)[j] = make_double4(s[0], s[1], s[2], s[3]);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v) u[j * V + v] = s[v];
    }
}

struct EqParams {
    double a1, a2, gamma;
};

// Physical fluxes in x and y: Trixi's flux(u, 1|2, equations) for the
// linear advection and compressible Euler equations (the latter evaluated
// with fused multiply-adds like Trixi's @muladd code).
template <int V>
__device__ __forceinline__ void flux_xy(const EqParams& e, const double* u, double* f1,
                                        double* f2)
{
    if constexpr (V == 1) {
        f1[0] = e.a1 * u[0];
        f2[0] = e.a2 * u[0];
    } else {
        const double rho = u[0], m1 = u[1], m2 = u[2], E = u[3];
        const double v1 = m1 / rho, v2 = m2 / rho;
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
    }
}

// Boundary/initial value functions (see MF_FN_* in mfrbf.h).
template <int V>
__device__ __forceinline__ void bc_value(int fn, const double* q, double gamma, double x,
                                         double y, double t, double* out)
{
    if (fn == MF_FN_CONST) {
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = q[v];
    } else if (fn == MF_FN_PULSE) {
        const double dx = (x - q[3] * t) - q[0], dy = (y - q[4] * t) - q[1];
        out[0] = exp(-(dx * dx + dy * dy) / (q[2] * q[2]));
#pragma unroll
        for (int v = 1; v < V; ++v) out[v] = 0.0;
    } else if (fn == MF_FN_WAVE) {
        if constexpr (V == 4) {
            const double arg = (x + y) - t * (q[2] + q[3]);
            const double rho = q[0] + q[1] * sin(2.0 * M_PI * arg + q[5]);
            out[0] = rho;
            out[1] = rho * q[2];
            out[2] = rho * q[3];
            out[3] = q[4] / (gamma - 1.0) + 0.5 * rho * (q[2] * q[2] + q[3] * q[3]);
        }
    } else if (fn == MF_FN_GRADIENT) {
        if constexpr (V == 4) {
            out[0] = 1.4 + 0.1 * x + 0.1 * y;
            out[1] = 4.1 + 0.1 * x + 0.1 * y;
            out[2] = 0.0 + 0.1 * x + 0.1 * y;
            out[3] = 8.8 + 0.1 * x + 0.1 * y;
        }
    }
}

// Julia's max(x, y) for floats: NaN wins, otherwise the larger value.
__device__ __forceinline__ double jmax(double x, double y)
{
    if (isnan(x)) return x;
    if (isnan(y)) return y;
    return (y > x || (signbit(y) < signbit(x))) ? y : x;
}

// Julia isless over 4-vectors: the first differing component decides,
// NaN above every number.
__device__ __forceinline__ bool lex_less(const double* a, const double* b)
{
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const bool an = isnan(a[v]), bn = isnan(b[v]);
        if (an && bn) continue;
        if (bn) return true;
        if (an) return false;
        if (a[v] < b[v]) return true;
        if (a[v] > b[v]) return false;
    }
    return false;
}

// Fixed-order block reductions over 4-vectors held in LDS (kBlock entries).
// Block-wide sum of every thread's 4 values: butterfly within each
// wavefront (a + b and b + a round alike, so all lanes hold the same bits),
// then the 4 wavefront sums in a fixed order.  Every thread gets the result.
__device__ __forceinline__ void block_sum4(double* v)
{
    __shared__ double wres[kBlock / 64][4];
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] += __shfl_xor(v[q], o);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) wres[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double t = wres[0][q];
        for (int w = 1; w < kBlock / 64; ++w) t = t + wres[w][q];
        v[q] = t;
    }
    __syncthreads();
}

// Block-wide lexicographic maximum of every thread's 4-vector, the same way.
__device__ __forceinline__ void block_lexmax4(double* v)
{
    __shared__ double wres[kBlock / 64][4];
    for (int o = 32; o > 0; o >>= 1) {
        double c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = __shfl_xor(v[q], o);
        if (lex_less(v, c))
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = c[q];
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) wres[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = wres[0][q];
    for (int w = 1; w < kBlock / 64; ++w)
        if (lex_less(v, wres[w]))
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = wres[w][q];
    __syncthreads();
}

// mean(u) from the per-block partial sums (every caller gets the same bits).
__device__ __forceinline__ void mean_from_partials(const double* psum, int nparts, int64_t N, double* mean)
{
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += kBlock)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] += psum[4 * b + v];
    block_sum4(acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) mean[v] = acc[v] / (double)(4 * N);
}

// ninf from the per-block partial lexicographic maxima, zeros -> eps().
__device__ __forceinline__ void ninf_from_partials(const double* pmax, int nparts, double* ninf)
{
    double best[4] = {-1.0, -1.0, -1.0, -1.0};
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
        double c[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) c[v] = pmax[4 * b + v];
        if (lex_less(best, c))
#pragma unroll
            for (int v = 0; v < 4; ++v) best[v] = c[v];
    }
    block_lexmax4(best);
#pragma unroll
    for (int v = 0; v < 4; ++v) ninf[v] = (best[v] == 0.0) ? DBL_EPSILON : best[v];
}

// XCD-aware logical block (the remap of remapped_node).
__device__ __forceinline__ int64_t remapped_block()
{
    const unsigned b = blockIdx.x, per = gridDim.x >> 3;
    return (int64_t)((b & 7u) * per + (b >> 3));
}

// LDS records are read by the 64 lanes of a wavefront at scattered record
// indices, one 16-byte chunk per lane per instruction.  With a plain stride
// of NC 16-byte chunks per record (NC = 2 or 4) those reads reach only a few
// LDS banks; a record's chunks are therefore stored in an order rotated by
// bits 2-3 of its index, which spreads one instruction's lanes over all the
// banks.  Layout only: every value and operation order is unchanged.
constexpr bool kLdsSwizzle = true;

template <int NC>
__device__ __forceinline__ int64_t lds_chunk(int64_t q, int c)
{
    if constexpr (kLdsSwizzle && NC == 4)
        return 4 * q + (c ^ (int)((q >> 2) & 3));
    else if constexpr (kLdsSwizzle && NC == 2)
        return 2 * q + (c ^ (int)((q >> 2) & 1));
    else
        return (int64_t)NC * q + c;
}

// Unit u (4 values) of LDS record q, records of RU units: FP64 (two chunks
// per unit) and FP32 (one chunk per unit).
template <int RU>
__device__ __forceinline__ void lds_store_unit(double* lds, int64_t q, int u, const double* s)
{
    double2* p = reinterpret_cast<double2*>(lds);
    p[lds_chunk<2 * RU>(q, 2 * u)] = make_double2(s[0], s[1]);
    p[lds_chunk<2 * RU>(q, 2 * u + 1)] = make_double2(s[2], s[3]);
}

template <int RU>
__device__ __forceinline__ void lds_load_unit(const double* lds, int64_t q, int u, double* s)
{
    const double2* p = reinterpret_cast<const double2*>(lds);
    const double2 a = p[lds_chunk<2 * RU>(q, 2 * u)], b = p[lds_chunk<2 * RU>(q, 2 * u + 1)];
    s[0] = a.x;
    s[1] = a.y;
    s[2] = b.x;
    s[3] = b.y;
}

template <int RU>
__device__ __forceinline__ void lds_store_unit32(float* lds, int64_t q, int u, const double* s)
{
    reinterpret_cast<float4*>(lds)[lds_chunk<RU>(q, u)] = make_float4((float)s[0], (float)s[1], (float)s[2],
                                                                      (float)s[3]);
}

template <int RU>
__device__ __forceinline__ void lds_load_unit32(const float* lds, int64_t q, int u, double* s)
{
    const float4 a = reinterpret_cast<const float4*>(lds)[lds_chunk<RU>(q, u)];
    s[0] = a.x;
    s[1] = a.y;
    s[2] = a.z;
    s[3] = a.w;
}

// ------------------------------------------------ RBF-FD weight helpers
// shared by the Bunch-Kaufman weights (setup.hip) and the null-space weights
// (weights_ns.hip)

__device__ __forceinline__ double ipow(double v, int e)
{
    double r = 1.0;
    for (int i = 0; i < e; ++i) r = r * v;
    return r;
}

__device__ __forceinline__ void mono_exp(int t, int* ax, int* ay)
{
    int d = 0;
    while (t >= d + 1) {
        t -= d + 1;
        ++d;
    }
    *ax = d - t;
    *ay = t;
}

__device__ __forceinline__ double phs_eval(const PhsTable& tb, double s, double x, double y)
{
#pragma clang fp contract(off)
    const double r = sqrt(x * x + y * y);
    double val = 0.0;
    for (int t = 0; t < tb.nt; ++t) {
        if (tb.c[t] == 0.0) continue;
        double term = tb.c[t] * ipow(s, tb.a[t]);
        if (tb.b[t] >= 0)
            term = term * ipow(r, tb.b[t]);
        else
            term = term / ipow(r, -tb.b[t]);
        val = val + term;
    }
    return val;
}

// s^k nearly correctly rounded (double-double product).
__device__ __forceinline__ double pow_k(double s, int k)
{
    double hi = 1.0, lo = 0.0;
    for (int r = 0; r < k; ++r) {
        const double ph = hi * s;
        const double pl = fma(hi, s, -ph) + lo * s;
        hi = ph + pl;
        lo = pl - (hi - ph);
    }
    return hi;
}

}  // namespace mf
