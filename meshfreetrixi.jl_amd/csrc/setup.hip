// Setup kernels: kNN stencils, RBF-FD weights, operator packing and the
// transposed operator used by the viscosity sources.
#include <hip/hip_runtime.h>

#include <vector>

#include <algorithm>
#include <cfloat>
#include <hipcub/hipcub.hpp>
#include <mutex>

#include "mf_device.h"

namespace mf {

// ================================================================= kNN
// Reference: PointData(points, basis), geometry_primatives.jl:322-339
// (NearestNeighbors knn(KDTree(points), points, nv, true)).  Uniform grid
// binning; one thread per query keeps its k best (d2, index) pairs in LDS
// ordered by squared distance, ties by index; rings of cells are searched
// until the k-th candidate is closer than every unvisited cell.
// d2 = (xi - xj)^2 + (yi - yj)^2 without fused multiply-add, as the
// reference's distance evaluation.

__global__ void k_bbox(int64_t N, const double* x, const double* y, double* out)
{
    __shared__ double r[4][1024];
    double a = INFINITY, b = -INFINITY, c = INFINITY, d = -INFINITY;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        a = fmin(a, x[i]);
        b = fmax(b, x[i]);
        c = fmin(c, y[i]);
        d = fmax(d, y[i]);
    }
    r[0][threadIdx.x] = a;
    r[1][threadIdx.x] = b;
    r[2][threadIdx.x] = c;
    r[3][threadIdx.x] = d;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r[0][threadIdx.x] = fmin(r[0][threadIdx.x], r[0][threadIdx.x + w]);
            r[1][threadIdx.x] = fmax(r[1][threadIdx.x], r[1][threadIdx.x + w]);
            r[2][threadIdx.x] = fmin(r[2][threadIdx.x], r[2][threadIdx.x + w]);
            r[3][threadIdx.x] = fmax(r[3][threadIdx.x], r[3][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) out[q] = r[q][0];
}

// ========================================================== node order
// Internal node order: the Morton (Z-order) code of each position,
// quantised to 16 bits per axis over the bounding square, ties kept in the
// caller's order (the radix sort is stable).  Stencil neighbours then sit
// close together in memory, so a block's neighbour gathers touch few cache
// lines.  The order changes no result: kNN ties and the reference-order
// sorts compare the caller's indices.
__device__ __forceinline__ uint32_t spread16(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void k_morton(int64_t N, const double* x, const double* y, const double* box, uint32_t* code,
                         int32_t* iota)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double ext = fmax(box[1] - box[0], box[3] - box[2]);
    if (!(ext > 0.0)) ext = 1.0;
    const double fx = (x[i] - box[0]) / ext * 65536.0, fy = (y[i] - box[2]) / ext * 65536.0;
    const uint32_t qx = (uint32_t)fmin(fmax(fx, 0.0), 65535.0);
    const uint32_t qy = (uint32_t)fmin(fmax(fy, 0.0), 65535.0);
    code[i] = spread16(qx) | (spread16(qy) << 1);
    iota[i] = (int32_t)i;
}

// Group-major Morton keys: (group << 32) | morton code of the point.
__global__ void k_morton_grouped(int64_t N, const double* x, const double* y, const double* box,
                                 const int32_t* group, uint64_t* code, int32_t* iota)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double ext = fmax(box[1] - box[0], box[3] - box[2]);
    if (!(ext > 0.0)) ext = 1.0;
    const double fx = (x[i] - box[0]) / ext * 65536.0, fy = (y[i] - box[2]) / ext * 65536.0;
    const uint32_t qx = (uint32_t)fmin(fmax(fx, 0.0), 65535.0);
    const uint32_t qy = (uint32_t)fmin(fmax(fy, 0.0), 65535.0);
    code[i] = ((uint64_t)(uint32_t)group[i] << 32) | (uint64_t)(spread16(qx) | (spread16(qy) << 1));
    iota[i] = (int32_t)i;
}

__global__ void k_iota(int64_t N, int32_t* v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) v[i] = (int32_t)i;
}

__global__ void k_apply_order(int64_t N, const int32_t* uid, const double* xu, const double* yu, double* x,
                              double* y, int32_t* inv)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int32_t u = uid[p];
    x[p] = xu[u];
    y[p] = yu[u];
    inv[u] = (int32_t)p;
}

int launch_node_order(mf_ctx* c, const double* d_xu, const double* d_yu, const int32_t* d_group)
{
    const int64_t N = c->N;
    const unsigned g = (unsigned)((N + 255) / 256);
    int rc;
    if (c->node_order != 1 && d_group) {
        double* box = nullptr;
        uint64_t *code = nullptr, *code_s = nullptr;
        int32_t* iota = nullptr;
        if ((rc = dalloc(c, &box, 4)) || (rc = dalloc(c, &code, N)) || (rc = dalloc(c, &code_s, N)) ||
            (rc = dalloc(c, &iota, N)))
            return rc;
        hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, c->stream, N, d_xu, d_yu, box);
        hipLaunchKernelGGL(k_morton_grouped, dim3(g), dim3(256), 0, c->stream, N, d_xu, d_yu, box, d_group, code,
                           iota);
        size_t tb = 0;
        MF_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, code, code_s, iota, c->d_uid, (int)N, 0, 64,
                                                     c->stream));
        void* tmp = nullptr;
        MF_HIP(c, hipMalloc(&tmp, tb + 16));
        MF_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, code, code_s, iota, c->d_uid, (int)N, 0, 64,
                                                     c->stream));
        MF_HIP(c, hipStreamSynchronize(c->stream));
        (void)hipFree(tmp);
        dfree(box);
        dfree(code);
        dfree(code_s);
        dfree(iota);
    } else if (c->node_order == 1) {
        hipLaunchKernelGGL(k_iota, dim3(g), dim3(256), 0, c->stream, N, c->d_uid);
    } else {
        double* box = nullptr;
        int32_t *code = nullptr, *code_s = nullptr, *iota = nullptr;
        if ((rc = dalloc(c, &box, 4)) || (rc = dalloc(c, &code, N)) || (rc = dalloc(c, &code_s, N)) ||
            (rc = dalloc(c, &iota, N)))
            return rc;
        uint32_t* ucode = reinterpret_cast<uint32_t*>(code);
        uint32_t* ucode_s = reinterpret_cast<uint32_t*>(code_s);
        hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, c->stream, N, d_xu, d_yu, box);
        hipLaunchKernelGGL(k_morton, dim3(g), dim3(256), 0, c->stream, N, d_xu, d_yu, box, ucode, iota);
        size_t tb = 0;
        MF_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ucode, ucode_s, iota, c->d_uid, (int)N, 0, 32,
                                                     c->stream));
        void* tmp = nullptr;
        MF_HIP(c, hipMalloc(&tmp, tb + 16));
        MF_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, ucode, ucode_s, iota, c->d_uid, (int)N, 0, 32,
                                                     c->stream));
        MF_HIP(c, hipStreamSynchronize(c->stream));
        (void)hipFree(tmp);
        dfree(box);
        dfree(code);
        dfree(code_s);
        dfree(iota);
    }
    hipLaunchKernelGGL(k_apply_order, dim3(g), dim3(256), 0, c->stream, N, c->d_uid, d_xu, d_yu, c->d_x, c->d_y,
                       c->d_inv);
    MF_HIP(c, hipGetLastError());
    c->h_inv.resize(N);
    MF_HIP(c, hipMemcpyAsync(c->h_inv.data(), c->d_inv, sizeof(int32_t) * N, hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

struct Grid {
    double x0, y0, cs;
    int64_t ncx, ncy;
};

__device__ __forceinline__ int64_t cell_of(const Grid& g, double px, double py, int64_t* cx,
                                           int64_t* cy)
{
    int64_t ix = (int64_t)floor((px - g.x0) / g.cs), iy = (int64_t)floor((py - g.y0) / g.cs);
    ix = ix < 0 ? 0 : (ix >= g.ncx ? g.ncx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= g.ncy ? g.ncy - 1 : iy);
    *cx = ix;
    *cy = iy;
    return iy * g.ncx + ix;
}

__global__ void k_cell_count(int64_t N, const double* x, const double* y, Grid g, int32_t* cell,
                             int32_t* count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int64_t cx, cy;
    const int64_t c = cell_of(g, x[i], y[i], &cx, &cy);
    cell[i] = (int32_t)c;
    atomicAdd(&count[c], 1);
}

__global__ void k_cell_scatter(int64_t N, const int32_t* cell, const int32_t* start,
                               int32_t* cursor, int32_t* order)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int32_t c = cell[i];
    const int32_t pos = start[c] + atomicAdd(&cursor[c], 1);
    order[pos] = (int32_t)i;
}

// Sort each cell's members by index (cells hold a handful of points) and
// gather their coordinates in cell order.
__global__ void k_cell_finish(int64_t ncell, const int32_t* start, const int32_t* count,
                              int32_t* order, const double* x, const double* y, double2* xys)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int32_t s = start[c], n = count[c];
    for (int a = 1; a < n; ++a) {
        const int32_t v = order[s + a];
        int b = a - 1;
        while (b >= 0 && order[s + b] > v) {
            order[s + b + 1] = order[s + b];
            --b;
        }
        order[s + b + 1] = v;
    }
    for (int a = 0; a < n; ++a) {
        const int32_t j = order[s + a];
        xys[s + a] = make_double2(x[j], y[j]);
    }
}

// The search grid over the bounding box: square cells of about two points
// each, their side at least the longer extent / N, so that an elongated or
// collinear cloud gets at most N + 1 cells along its length instead of a
// count that overflows; then W H / cs^2 <= N / 2 and the grid has at most
// 2.5 N + 1 cells.  The search is exact on any grid; only its speed
// depends on the cells.
static Grid knn_grid(const double box[4], int64_t N)
{
    Grid g;
    g.x0 = box[0];
    g.y0 = box[2];
    const double W = fmax(box[1] - box[0], 0.0), H = fmax(box[3] - box[2], 0.0);
    double cs = sqrt(2.0 * W * H / (double)N);
    cs = fmax(cs, fmax(W, H) / (double)N);
    if (!(cs > 0)) cs = 1.0;  // every point at one place: one cell
    g.cs = cs;
    g.ncx = (int64_t)fmin(floor(W / cs), (double)N) + 1;
    g.ncy = (int64_t)fmin(floor(H / cs), (double)N) + 1;
    return g;
}

__device__ __forceinline__ bool knn_before(double da, int32_t ia, double db, int32_t ib)
{
    return (da < db) || (da == db && ia < ib);
}

struct KnnArgs {
    int64_t N;
    int k;
    Grid g;
    const int32_t* start;
    const int32_t* count;
    const int32_t* order;
    const double2* xys;
    const int32_t* uid;  // internal -> caller index (ties compare caller indices)
    const int32_t* inv;  // caller -> internal index
    int32_t* out_idx;    // slot-major, internal indices (rows = 0)
    double* out_d2;      // slot-major squared distances (rows = 0)
    int rows;            // 1: out_idx row per caller node with caller indices,
                         //    out_d2 the distances (not squared), same layout
};

__global__ __launch_bounds__(64) void k_knn_query(const KnnArgs a)
{
#pragma clang fp contract(off)
    extern __shared__ char smem[];
    const int lane = threadIdx.x;
    double* bd = reinterpret_cast<double*>(smem);
    int32_t* bi = reinterpret_cast<int32_t*>(smem + sizeof(double) * 64 * a.k);
    const int64_t p = (int64_t)blockIdx.x * 64 + lane;  // position in cell order
    if (p >= a.N) return;
    const int32_t q = a.order[p];
    const double2 qp = a.xys[p];
    const double qx = qp.x, qy = qp.y;
    int64_t cx, cy;
    cell_of(a.g, qx, qy, &cx, &cy);
    const int K = a.k;
    int cnt = 0;
    for (int64_t r = 0;; ++r) {
        const int64_t ax = cx - r, bx = cx + r, ay = cy - r, by = cy + r;
        for (int64_t jy = ay; jy <= by; ++jy) {
            if (jy < 0 || jy >= a.g.ncy) continue;
            for (int64_t jx = ax; jx <= bx; ++jx) {
                if (jx < 0 || jx >= a.g.ncx) continue;
                if (r > 0 && jy != ay && jy != by && jx != ax && jx != bx) continue;
                const int64_t c = jy * a.g.ncx + jx;
                const int32_t s0 = a.start[c], s1 = s0 + a.count[c];
                for (int32_t pp = s0; pp < s1; ++pp) {
                    const double2 cp = a.xys[pp];
                    const double dx = qx - cp.x, dy = qy - cp.y;
                    const double d2 = dx * dx + dy * dy;
                    const int32_t j = a.uid[a.order[pp]];
                    if (cnt == K && !knn_before(d2, j, bd[(K - 1) * 64 + lane], bi[(K - 1) * 64 + lane]))
                        continue;
                    int pos = (cnt < K) ? cnt : K - 1;
                    while (pos > 0 && knn_before(d2, j, bd[(pos - 1) * 64 + lane], bi[(pos - 1) * 64 + lane])) {
                        bd[pos * 64 + lane] = bd[(pos - 1) * 64 + lane];
                        bi[pos * 64 + lane] = bi[(pos - 1) * 64 + lane];
                        --pos;
                    }
                    bd[pos * 64 + lane] = d2;
                    bi[pos * 64 + lane] = j;
                    if (cnt < K) ++cnt;
                }
            }
        }
        const bool all = (ax <= 0 && bx >= a.g.ncx - 1 && ay <= 0 && by >= a.g.ncy - 1);
        if (all) break;
        if (cnt == K) {
            double dm = INFINITY;
            if (ax > 0) dm = fmin(dm, qx - (a.g.x0 + (double)ax * a.g.cs));
            if (bx < a.g.ncx - 1) dm = fmin(dm, (a.g.x0 + (double)(bx + 1) * a.g.cs) - qx);
            if (ay > 0) dm = fmin(dm, qy - (a.g.y0 + (double)ay * a.g.cs));
            if (by < a.g.ncy - 1) dm = fmin(dm, (a.g.y0 + (double)(by + 1) * a.g.cs) - qy);
            dm *= (1.0 - 1e-9);
            if (dm > 0 && bd[(K - 1) * 64 + lane] < dm * dm) break;
        }
    }
    if (a.rows) {
        const int64_t row = (int64_t)a.uid[q] * K;
        for (int s = 0; s < K; ++s) {
            a.out_idx[row + s] = bi[s * 64 + lane];
            if (a.out_d2) a.out_d2[row + s] = sqrt(bd[s * 64 + lane]);
        }
        return;
    }
    for (int s = 0; s < K; ++s) {
        a.out_idx[(int64_t)s * a.N + q] = a.inv[bi[s * 64 + lane]];
        if (a.out_d2) a.out_d2[(int64_t)s * a.N + q] = bd[s * 64 + lane];
    }
}

int launch_knn(mf_ctx* c, int k, double* d_d2)
{
    const int64_t N = c->N;
    double* d_box = nullptr;
    int rc = dalloc(c, &d_box, 4);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, c->stream, N, c->d_x, c->d_y, d_box);
    double box[4];
    MF_HIP(c, hipMemcpyAsync(box, d_box, sizeof box, hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(d_box);
    const Grid g = knn_grid(box, N);
    const int64_t ncell = g.ncx * g.ncy;
    if (ncell > (int64_t)1 << 30) return fail(c, MF_ERR_ARGUMENT, "kNN: search grid too large");
    int32_t *cell = nullptr, *count = nullptr, *start = nullptr, *cursor = nullptr, *order = nullptr;
    double2* xys = nullptr;
    if ((rc = dalloc(c, &cell, N)) || (rc = dalloc(c, &count, ncell)) || (rc = dalloc(c, &start, ncell)) ||
        (rc = dalloc(c, &cursor, ncell)) || (rc = dalloc(c, &order, N)) || (rc = dalloc(c, &xys, N)))
        return rc;
    MF_HIP(c, hipMemsetAsync(count, 0, sizeof(int32_t) * ncell, c->stream));
    MF_HIP(c, hipMemsetAsync(cursor, 0, sizeof(int32_t) * ncell, c->stream));
    const unsigned gN = (unsigned)((N + 255) / 256), gC = (unsigned)((ncell + 255) / 256);
    hipLaunchKernelGGL(k_cell_count, dim3(gN), dim3(256), 0, c->stream, N, c->d_x, c->d_y, g, cell, count);
    size_t tmp_bytes = 0;
    MF_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, count, start, (int)ncell, c->stream));
    void* tmp = nullptr;
    MF_HIP(c, hipMalloc(&tmp, tmp_bytes + 16));
    MF_HIP(c, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, count, start, (int)ncell, c->stream));
    hipLaunchKernelGGL(k_cell_scatter, dim3(gN), dim3(256), 0, c->stream, N, cell, start, cursor, order);
    hipLaunchKernelGGL(k_cell_finish, dim3(gC), dim3(256), 0, c->stream, ncell, start, count, order,
                       c->d_x, c->d_y, xys);
    KnnArgs a;
    a.N = N;
    a.k = k;
    a.g = g;
    a.start = start;
    a.count = count;
    a.order = order;
    a.xys = xys;
    a.uid = c->d_uid;
    a.inv = c->d_inv;
    a.out_idx = c->d_idx;
    a.out_d2 = d_d2;
    a.rows = 0;
    const size_t lds = (size_t)64 * k * (sizeof(double) + sizeof(int32_t));
    hipLaunchKernelGGL(k_knn_query, dim3((unsigned)((N + 63) / 64)), dim3(64), lds, c->stream, a);
    MF_HIP(c, hipGetLastError());
    MF_HIP(c, hipStreamSynchronize(c->stream));
    (void)hipFree(tmp);
    dfree(cell);
    dfree(count);
    dfree(start);
    dfree(cursor);
    dfree(order);
    dfree(xys);
    return MF_OK;
}

// ------------------------------------------------ kNN without allocation
// mf_knn_noalloc runs the same grid search with every intermediate (the
// bounding box, the cell of each point, the cell counts, starts and fill
// cursors, the points in cell order and their coordinates) in `scratch`, a
// file-scope device array of fixed size allocated with the code object:
// up to kKnnScratchNodes points and as many grid cells.  Host code uses
// the array only through the address hipGetSymbolAddress returns for the
// current device (knn_scratch_ptr); every call on a device shares its
// copy, so calls that overlap on one device race.
constexpr int64_t kKnnScratchNodes = 262144;
constexpr int64_t kKnnScratchCells = 262144;
constexpr size_t kKnnScratchBytes = 256 + 3 * kKnnScratchCells * sizeof(int32_t) +
                                    kKnnScratchNodes * (2 * sizeof(int32_t) + sizeof(double2));
static __device__ __align__(256) unsigned char scratch[kKnnScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_knn_scratch[kMaxDevices];
static std::mutex g_knn_scratch_mu;

int64_t knn_noalloc_max_nodes() { return kKnnScratchNodes; }

// Device address of `scratch` on the current device (looked up once per device).
static int knn_scratch_ptr(mf_ctx* c, unsigned char** out)
{
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, MF_ERR_ARGUMENT, "kNN scratch: device index");
    std::lock_guard<std::mutex> lock(g_knn_scratch_mu);
    if (!g_knn_scratch[c->device]) {
        void* ptr = nullptr;
        MF_HIP(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(c, MF_ERR_HIP, "kNN scratch: null symbol address");
        g_knn_scratch[c->device] = ptr;
    }
    *out = static_cast<unsigned char*>(g_knn_scratch[c->device]);
    return MF_OK;
}

__global__ void k_fill_i32(int64_t n, int32_t* p, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Exclusive prefix sum of n counts in one workgroup of 1024 threads: each
// thread sums a contiguous chunk, the chunk totals are scanned in LDS, and
// each thread writes its chunk's prefixes.
__global__ __launch_bounds__(1024) void k_scan_excl(int64_t n, const int32_t* in, int32_t* out)
{
    __shared__ int32_t tot[1024];
    const int t = threadIdx.x;
    const int64_t chunk = (n + 1023) / 1024;
    const int64_t lo = (int64_t)t * chunk, hi = lo + chunk < n ? lo + chunk : n;
    int32_t sum = 0;
    for (int64_t i = lo; i < hi; ++i) sum += in[i];
    tot[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int32_t v = t >= o ? tot[t - o] : 0;
        __syncthreads();
        tot[t] += v;
        __syncthreads();
    }
    int32_t run = tot[t] - sum;  // exclusive prefix of this chunk
    for (int64_t i = lo; i < hi; ++i) {
        out[i] = run;
        run += in[i];
    }
}

int launch_knn_noalloc(mf_ctx* c, int k, int32_t* d_idx, double* d_dist)
{
    const int64_t N = c->N;
    unsigned char* base = nullptr;
    int rc = knn_scratch_ptr(c, &base);
    if (rc) return rc;
    double* d_box = reinterpret_cast<double*>(base);
    int32_t* count = reinterpret_cast<int32_t*>(base + 256);
    int32_t* start = count + kKnnScratchCells;
    int32_t* cursor = start + kKnnScratchCells;
    int32_t* cell = cursor + kKnnScratchCells;
    int32_t* order = cell + kKnnScratchNodes;
    double2* xys = reinterpret_cast<double2*>(order + kKnnScratchNodes);
    hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, c->stream, N, c->d_x, c->d_y, d_box);
    double box[4];
    MF_HIP(c, hipMemcpyAsync(box, d_box, sizeof box, hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    const Grid g = knn_grid(box, N);
    const int64_t ncell = g.ncx * g.ncy;
    if (ncell > kKnnScratchCells)
        return fail(c, MF_ERR_UNSUPPORTED, "mf_knn_noalloc: the cloud's search grid needs more cells than its "
                                           "static scratch holds (262,144; an elongated cloud)");
    const unsigned gN = (unsigned)((N + 255) / 256), gC = (unsigned)((ncell + 255) / 256);
    hipLaunchKernelGGL(k_fill_i32, dim3(gC), dim3(256), 0, c->stream, ncell, count, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(gC), dim3(256), 0, c->stream, ncell, cursor, 0);
    hipLaunchKernelGGL(k_cell_count, dim3(gN), dim3(256), 0, c->stream, N, c->d_x, c->d_y, g, cell, count);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, c->stream, ncell, count, start);
    hipLaunchKernelGGL(k_cell_scatter, dim3(gN), dim3(256), 0, c->stream, N, cell, start, cursor, order);
    hipLaunchKernelGGL(k_cell_finish, dim3(gC), dim3(256), 0, c->stream, ncell, start, count, order,
                       c->d_x, c->d_y, xys);
    KnnArgs a;
    a.N = N;
    a.k = k;
    a.g = g;
    a.start = start;
    a.count = count;
    a.order = order;
    a.xys = xys;
    a.uid = c->d_uid;
    a.inv = c->d_inv;
    a.out_idx = d_idx;
    a.out_d2 = d_dist;
    a.rows = 1;
    const size_t lds = (size_t)64 * k * (sizeof(double) + sizeof(int32_t));
    hipLaunchKernelGGL(k_knn_query, dim3((unsigned)((N + 63) / 64)), dim3(64), lds, c->stream, a);
    MF_HIP(c, hipGetLastError());
    MF_HIP(c, hipStreamSynchronize(c->stream));
    return MF_OK;
}

// dx_min = min, dx_avg = mean of the nearest-neighbour distance
// (geometry_primatives.jl:332-334).  The mean sums in the caller's node
// order exactly as Julia's `sum` does (Base.mapreduce_impl: halve the
// range until it has at most 1024 entries, add each such block left to
// right, then add the halves back up), so dx_avg, which scales the upwind
// and residual viscosities, has the reference's bits: the GPU forms every
// block's sum (and minimum), the host adds the blocks along the same
// recursion.
__global__ void k_dx_blocks(int64_t nb, const int64_t* start, const double* d2nn, const int32_t* inv,
                            double* bsum, double* bmin)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int64_t lo = start[b], hi = start[b + 1];
    double v = sqrt(d2nn[inv[lo]]), mn = v;
    for (int64_t i = lo + 1; i < hi; ++i) {
        const double d = sqrt(d2nn[inv[i]]);
        v = v + d;
        mn = fmin(mn, d);
    }
    bsum[b] = v;
    bmin[b] = mn;
}

// The blocks of the pairwise recursion over [lo, hi], in order.
static void pairwise_blocks(int64_t lo, int64_t hi, std::vector<int64_t>& start)
{
    if (hi - lo < 1024) {
        start.push_back(lo);
        return;
    }
    const int64_t mid = lo + ((hi - lo) >> 1);
    pairwise_blocks(lo, mid, start);
    pairwise_blocks(mid + 1, hi, start);
}

static double pairwise_combine(int64_t lo, int64_t hi, const std::vector<double>& bsum, size_t& next)
{
    if (hi - lo < 1024) return bsum[next++];
    const int64_t mid = lo + ((hi - lo) >> 1);
    const double a = pairwise_combine(lo, mid, bsum, next);
    return a + pairwise_combine(mid + 1, hi, bsum, next);
}

int launch_dx_stats(mf_ctx* c, const double* d_d2, double* dx_min, double* dx_avg)
{
    const int64_t N = c->N;
    std::vector<int64_t> start;
    pairwise_blocks(0, N - 1, start);
    const int64_t nb = (int64_t)start.size();
    start.push_back(N);
    int64_t* d_start = nullptr;
    double* d_out = nullptr;
    int rc;
    if ((rc = dalloc(c, &d_start, start.size())) || (rc = dalloc(c, &d_out, 2 * (size_t)nb))) {
        dfree(d_start);
        return rc;
    }
    MF_HIP(c, hipMemcpyAsync(d_start, start.data(), sizeof(int64_t) * start.size(), hipMemcpyHostToDevice,
                             c->stream));
    hipLaunchKernelGGL(k_dx_blocks, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, c->stream, nb, d_start,
                       d_d2 + N, c->d_inv, d_out, d_out + nb);
    MF_HIP(c, hipGetLastError());
    std::vector<double> h(2 * (size_t)nb);
    MF_HIP(c, hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(d_start);
    dfree(d_out);
    std::vector<double> bsum(h.begin(), h.begin() + nb);
    double mn = h[nb];
    for (int64_t q = 1; q < nb; ++q) mn = std::min(mn, h[nb + q]);
    size_t next = 0;
    *dx_min = mn;
    *dx_avg = pairwise_combine(0, N - 1, bsum, next) / (double)N;
    return MF_OK;
}

// ============================================================= weights
// Reference: compute_flux_operator, compute_operators.jl:409-453 and
// :549-594.  One wavefront per stencil.  The m x m saddle matrix
// [R P; P' 0] lives in LDS (column-major, leading dimension ld); lane i
// owns row i while assembling.  The solve restates LAPACK DSYTF2/DSYTRS
// (UPLO='U', Bunch-Kaufman pivoting), which Julia's Symmetric(M) \ rhs
// (:435) calls; the operation order matches that algorithm so the pivot
// sequence is the reference's.
struct WeightsArgs {
    const double* x;
    const double* y;
    const int64_t* ptr;
    const int32_t* order;  // stencils of this launch (ragged batches), or null
    int64_t base;          // first entry of order for this launch
    const int32_t* ridx;
    const int32_t* ell;
    int ell_k;
    int64_t N;
    int p, np, kd, ld;
    double kfact;
    PhsTable tab;
    double* wx;
    double* wy;
    int* status;
};

// Maximum over the wavefront of one double per lane (an exact selection,
// so any combination order gives the same bits), returned to every lane:
// a DPP max-scan within each row of 16 lanes (row_shr 1, 2, 4, 8), then
// across rows (row_bcast 15, 31), which leaves the maximum in lane 63.
// Lanes the shifted source does not reach take -1, below every |x|; every
// lane of the wavefront must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_shift(double v)
{
    constexpr int kNeg1Hi = (int)0xbff00000;  // -1.0 = 0xbff00000'00000000
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(kNeg1Hi, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_shift<0x111, 0xf>(v));  // row_shr:1
    v = fmax(v, dpp_shift<0x112, 0xf>(v));  // row_shr:2
    v = fmax(v, dpp_shift<0x114, 0xf>(v));  // row_shr:4
    v = fmax(v, dpp_shift<0x118, 0xf>(v));  // row_shr:8
    v = fmax(v, dpp_shift<0x142, 0xa>(v));  // row_bcast:15, rows 1 and 3
    v = fmax(v, dpp_shift<0x143, 0xc>(v));  // row_bcast:31, rows 2 and 3
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Lane l's value of v, to every lane (l uniform).
__device__ __forceinline__ double lane_value(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Interchange of rows p and q of B held one row per lane (p, q uniform).
__device__ __forceinline__ void lanes_swap(double& b0, double& b1, int lane, int p, int q)
{
    const double p0 = lane_value(b0, p), p1 = lane_value(b1, p);
    const double q0 = lane_value(b0, q), q1 = lane_value(b1, q);
    if (lane == p) {
        b0 = q0;
        b1 = q1;
    } else if (lane == q) {
        b0 = p0;
        b1 = p1;
    }
}

// sum_{i < n} b[i] x[i], added strictly in order i = 0, 1, ... (DGEMV 'T').
__device__ __forceinline__ double dot_rows(const double* b, const double* x, int n)
{
    double t = 0.0;
#pragma unroll 8
    for (int i = 0; i < n; ++i) t = t + b[i] * x[i];
    return t;
}

// The rank-1 and rank-2 updates of the factorization spread the flattened
// upper triangle (packed column by column, so element f of the triangle is
// A[f]) over the wavefront: lane l takes f = l + 64 t.  Its row and column
// for each t are the same at every pivot step, so they are decoded once per
// stencil into a register table, two 8-bit (row, column) pairs per word.
// Each element gets exactly the reference's arithmetic, so the spread
// changes no bits.
constexpr int kTriSlots = (kMaxStencilM * (kMaxStencilM - 1) / 2 + 63) / 64;  // t per lane

__device__ __forceinline__ void tri_table(int lane, uint32_t (&dec)[(kTriSlots + 1) / 2])
{
    int i = lane, j = 0;  // position of f = lane
    while (i > j) {
        i -= j + 1;
        ++j;
    }
#pragma unroll
    for (int t = 0; t < kTriSlots; ++t) {
        const uint32_t pair = (uint32_t)i | ((uint32_t)j << 8);
        if (t & 1)
            dec[t >> 1] |= pair << 16;
        else
            dec[t >> 1] = pair;
        i += 64;  // f + 64
        while (i > j) {
            i -= j + 1;
            ++j;
        }
    }
}

__global__ __launch_bounds__(64) void k_weights(const WeightsArgs a)
{
#pragma clang fp contract(off)
    extern __shared__ double sm[];
    const int ld = a.ld;
    // the upper triangle of the saddle matrix, packed column by column
    // (DSYTF2/DSYTRS with UPLO = 'U' never touch the lower one)
    double* A = sm;
    double* B = A + (size_t)ld * (ld + 1) / 2;
    double* X = B + 2 * ld;
    int* ipiv = reinterpret_cast<int*>(X + 2 * ld);
    const int64_t e = a.order ? (int64_t)a.order[a.base + blockIdx.x] : (int64_t)blockIdx.x;
    const int lane = threadIdx.x;
    int64_t off;
    int n;
    if (a.ptr) {
        off = a.ptr[e];
        n = (int)(a.ptr[e + 1] - off);
    } else {
        off = e;
        n = a.ell_k;
    }
    auto sidx = [&](int j) -> int32_t { return a.ptr ? a.ridx[off + j] : a.ell[(int64_t)j * a.N + e]; };
    const int m = n + a.np;
#define AA(i, j) A[(((j)-1) * (j)) / 2 + (i)-1]
#define BB(i, c) B[((i)-1) + (size_t)(c) * ld]
    // shift_stencil (:225-246): centre on node 0, scale each dimension
    const int32_t c0 = sidx(0);
    double px = 0.0, py = 0.0;
    if (lane < n) {
        const int32_t q = sidx(lane);
        px = a.x[q] - a.x[c0];
        py = a.y[q] - a.y[c0];
    }
    double mx = fabs(px), my = fabs(py);
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        my = fmax(my, __shfl_xor(my, o));
    }
    const double sx = 1.0 / mx, sy = 1.0 / my;
    if (lane < n) {
        X[2 * lane] = px * sx;
        X[2 * lane + 1] = py * sy;
    }
    __syncthreads();
    // interpolation_block [R P; P' 0] and right-hand sides [L_x, L_y]
    if (lane < m) {
        const int i = lane;
        if (i < n) {
            const double xi = X[2 * i], yi = X[2 * i + 1];
            for (int j = i; j < n; ++j) {
                const double dx = xi - X[2 * j], dy = yi - X[2 * j + 1];
                AA(i + 1, j + 1) = ipow(sqrt(dx * dx + dy * dy), a.p);
            }
            for (int t = 0; t < a.np; ++t) {
                int ax, ay;
                mono_exp(t, &ax, &ay);
                AA(i + 1, n + t + 1) = ipow(xi, ax) * ipow(yi, ay);
            }
            double qx = -xi, qy = -yi;
            if (i == 0) qx = qy = DBL_EPSILON;
            B[i] = phs_eval(a.tab, qx, qx, qy);
            B[i + ld] = phs_eval(a.tab, qy, qx, qy);
        } else {
            const int t = i - n;
            int ax, ay;
            mono_exp(t, &ax, &ay);
            for (int j = i; j < m; ++j) AA(i + 1, j + 1) = 0.0;
            B[i] = (ax == a.kd && ay == 0) ? a.kfact : 0.0;
            B[i + ld] = (ay == a.kd && ax == 0) ? a.kfact : 0.0;
        }
    }
    __syncthreads();
    // ---- DSYTF2, UPLO = 'U'
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    int info = 0;
    int k = m;
    uint32_t dec[(kTriSlots + 1) / 2];
    tri_table(lane, dec);
    double* W = X;  // reused: 2 x ld scratch for the 2x2-pivot multipliers
    while (k >= 1) {
        // the table is opaque to the compiler from step to step, so it is
        // decoded where used instead of held decoded across the whole loop
#pragma unroll
        for (int q = 0; q < (kTriSlots + 1) / 2; ++q) asm volatile("" : "+v"(dec[q]));
        int kstep = 1;
        const double absakk = fabs(AA(k, k));
        double colmax = 0.0;
        int imax = 0;
        if (k > 1) {
            // IDAMAX: the largest magnitude, the first row holding it
            const double v = (lane < k - 1) ? fabs(AA(lane + 1, k)) : -1.0;
            colmax = wave_max(v);
            imax = __ffsll((unsigned long long)__ballot(v == colmax));
        }
        int kp;
        if (fmax(absakk, colmax) == 0.0 || isnan(absakk)) {
            if (info == 0) info = k;
            kp = k;
        } else {
            if (absakk >= alpha * colmax) {
                kp = k;
            } else {
                // largest off-diagonal magnitude in row/column imax of the leading k x k block
                const int jj = lane + 1;
                double v = -1.0;
                if (jj > imax && jj <= k) v = fabs(AA(imax, jj));
                else if (jj < imax) v = fabs(AA(jj, imax));
                const double rowmax = wave_max(v);
                if (absakk >= alpha * colmax * (colmax / rowmax)) {
                    kp = k;
                } else if (fabs(AA(imax, imax)) >= alpha * rowmax) {
                    kp = imax;
                } else {
                    kp = imax;
                    kstep = 2;
                }
            }
            const int kk = k - kstep + 1;
            if (kp != kk) {
                const int ii = lane + 1;
                if (ii <= kp - 1) {
                    const double t = AA(ii, kk);
                    AA(ii, kk) = AA(ii, kp);
                    AA(ii, kp) = t;
                }
                if (ii >= kp + 1 && ii <= kk - 1) {
                    const double t = AA(ii, kk);
                    AA(ii, kk) = AA(kp, ii);
                    AA(kp, ii) = t;
                }
                if (lane == 0) {
                    const double t = AA(kk, kk);
                    AA(kk, kk) = AA(kp, kp);
                    AA(kp, kp) = t;
                    if (kstep == 2) {
                        const double t2 = AA(k - 1, k);
                        AA(k - 1, k) = AA(kp, k);
                        AA(kp, k) = t2;
                    }
                }
            }
            __syncthreads();
            if (kstep == 1) {
                // A(1:k-1, 1:k-1) += x x' * (-1/d), x = A(1:k-1, k) (DSYR, upper)
                const double r1 = 1.0 / AA(k, k);
                const int T = (k - 1) * k / 2;  // also where column k starts
                const double* xk = A + T;
#pragma unroll
                for (int t = 0; t < kTriSlots; ++t) {
                    const int f = lane + 64 * t;
                    if (64 * t >= T) break;
                    if (f < T) {
                        const uint32_t d = dec[t >> 1] >> ((t & 1) * 16);
                        const int ti = d & 0xff, tj = (d >> 8) & 0xff;
                        // DSYR skips a zero multiplier: the element is stored back unchanged
                        const double xj = xk[tj], xi = xk[ti], aij = A[f];
                        const double temp = -r1 * xj;
                        const double upd = aij + xi * temp;
                        A[f] = xj != 0.0 ? upd : aij;
                    }
                }
                __syncthreads();
                const int i = lane + 1;
                if (i <= k - 1) AA(i, k) = r1 * AA(i, k);
            } else if (k > 2) {
                double d12 = AA(k - 1, k);
                const double d22 = AA(k - 1, k - 1) / d12;
                const double d11 = AA(k, k) / d12;
                const double t = 1.0 / (d11 * d22 - 1.0);
                d12 = t / d12;
                const int j = lane + 1;
                if (j <= k - 2) {
                    W[j - 1] = d12 * (d11 * AA(j, k - 1) - AA(j, k));   // wkm1
                    W[ld + j - 1] = d12 * (d22 * AA(j, k) - AA(j, k - 1));  // wk
                }
                __syncthreads();
                // A(1:k-2, 1:k-2) -= [A(:,k) A(:,k-1)] [wk wkm1]' (upper)
                const int T = (k - 2) * (k - 1) / 2;  // also where column k-1 starts
                const double* xkm1 = A + T;
                const double* xk = A + (k - 1) * k / 2;
#pragma unroll
                for (int t = 0; t < kTriSlots; ++t) {
                    const int f = lane + 64 * t;
                    if (64 * t >= T) break;
                    if (f < T) {
                        const uint32_t d = dec[t >> 1] >> ((t & 1) * 16);
                        const int ti = d & 0xff, tj = (d >> 8) & 0xff;
                        A[f] = A[f] - xk[ti] * W[ld + tj] - xkm1[ti] * W[tj];
                    }
                }
                __syncthreads();
                if (j <= k - 2) {
                    AA(j, k) = W[ld + j - 1];
                    AA(j, k - 1) = W[j - 1];
                }
            }
        }
        if (lane == 0) {
            if (kstep == 1) {
                ipiv[k - 1] = kp;
            } else {
                ipiv[k - 1] = -kp;
                ipiv[k - 2] = -kp;
            }
        }
        __syncthreads();
        k -= kstep;
    }
    if (info && lane == 0) atomicOr(a.status, 1);
    // ---- DSYTRS, UPLO = 'U', two right-hand sides
    // First loop (U D solve, k = m down to 1): lane i holds row i + 1 of B in
    // registers; the pivot rows reach every lane by readlane, interchanges
    // are selects between two lanes.
    double b0 = 0.0, b1 = 0.0;
    if (lane < m) {
        b0 = BB(lane + 1, 0);
        b1 = BB(lane + 1, 1);
    }
    k = m;
    while (k >= 1) {
        const int K = k - 1;
        const int pv = __builtin_amdgcn_readfirstlane(ipiv[K]);
        if (pv > 0) {
            if (pv != k) lanes_swap(b0, b1, lane, K, pv - 1);
            const double bk0 = lane_value(b0, K), bk1 = lane_value(b1, K);
            if (lane < K) {
                const double aik = AA(lane + 1, k);
                if (bk0 != 0.0) b0 = b0 + aik * (-1.0 * bk0);
                if (bk1 != 0.0) b1 = b1 + aik * (-1.0 * bk1);
            }
            const double r = 1.0 / AA(k, k);
            if (lane == K) {
                b0 = r * b0;
                b1 = r * b1;
            }
            k -= 1;
        } else {
            const int kp = -pv;
            if (kp != k - 1) lanes_swap(b0, b1, lane, K - 1, kp - 1);
            const double bk0 = lane_value(b0, K), bk1 = lane_value(b1, K);
            const double ck0 = lane_value(b0, K - 1), ck1 = lane_value(b1, K - 1);
            if (lane < K - 1) {
                const double aik = AA(lane + 1, k), aikm1 = AA(lane + 1, k - 1);
                if (bk0 != 0.0) b0 = b0 + aik * (-1.0 * bk0);
                if (bk1 != 0.0) b1 = b1 + aik * (-1.0 * bk1);
                if (ck0 != 0.0) b0 = b0 + aikm1 * (-1.0 * ck0);
                if (ck1 != 0.0) b1 = b1 + aikm1 * (-1.0 * ck1);
            }
            const double akm1k = AA(k - 1, k);
            const double akm1 = AA(k - 1, k - 1) / akm1k;
            const double ak = AA(k, k) / akm1k;
            const double denom = akm1 * ak - 1.0;
            if (lane == K - 1) {
                b0 = (ak * (ck0 / akm1k) - bk0 / akm1k) / denom;
                b1 = (ak * (ck1 / akm1k) - bk1 / akm1k) / denom;
            } else if (lane == K) {
                b0 = (akm1 * (bk0 / akm1k) - ck0 / akm1k) / denom;
                b1 = (akm1 * (bk1 / akm1k) - ck1 / akm1k) / denom;
            }
            k -= 2;
        }
    }
    if (lane < m) {
        BB(lane + 1, 0) = b0;
        BB(lane + 1, 1) = b1;
    }
    __syncthreads();
    // Second loop (U' solve, k = 1 up to m): B(k) -= A(1:k-1, k)' B(1:k-1)
    // (DGEMV 'T', summed in row order), then the interchange; lane c < 2
    // takes right-hand side c.
    k = 1;
    while (k <= m) {
        const bool one = ipiv[k - 1] > 0;
        if (lane < 2) {
            const double* bc = &BB(1, lane);
            const double nk = BB(k, lane) + -1.0 * dot_rows(bc, &AA(1, k), k - 1);
            double nk1 = 0.0;
            if (!one) nk1 = BB(k + 1, lane) + -1.0 * dot_rows(bc, &AA(1, k + 1), k - 1);
            BB(k, lane) = nk;
            if (!one) BB(k + 1, lane) = nk1;
            const int kp = one ? ipiv[k - 1] : -ipiv[k - 1];
            if (kp != k) {
                const double t = BB(k, lane);
                BB(k, lane) = BB(kp, lane);
                BB(kp, lane) = t;
            }
        }
        __syncthreads();
        k += one ? 1 : 2;
    }
#undef AA
#undef BB
    // scale back (:438-439, :579-580)
    if (lane < n) {
        const double fx = pow_k(sx, a.kd), fy = pow_k(sy, a.kd);
        const int64_t o = a.ptr ? off + lane : (int64_t)lane * a.N + e;
        a.wx[o] = fx * B[lane];
        a.wy[o] = fy * B[lane + ld];
    }
}

int launch_weights(mf_ctx* c, const int64_t* d_ptr, const int32_t* d_idx, int ell_k, int64_t n_stencils,
                   int maxn, int p, int deg, int kd, double* d_wx, double* d_wy, int* d_status,
                   const int32_t* d_order, int64_t base)
{
    WeightsArgs a;
    a.x = c->d_x;
    a.y = c->d_y;
    a.ptr = d_ptr;
    a.order = d_order;
    a.base = base;
    a.ridx = d_ptr ? d_idx : nullptr;
    a.ell = d_ptr ? nullptr : d_idx;
    a.ell_k = ell_k;
    a.N = c->N;
    a.p = p;
    a.np = (deg + 1) * (deg + 2) / 2;
    a.kd = kd;
    const int mmax = maxn + a.np;
    if (mmax > kMaxStencilM)
        return fail(c, MF_ERR_UNSUPPORTED, "weights: stencil + polynomial size exceeds 64");
    a.ld = mmax | 1;
    double kf = 1.0;
    for (int i = 2; i <= kd; ++i) kf *= i;
    a.kfact = kf;
    a.tab = phs_table(p, kd);
    a.wx = d_wx;
    a.wy = d_wy;
    a.status = d_status;
    const size_t lds = sizeof(double) * ((size_t)a.ld * (a.ld + 1) / 2 + 4 * a.ld) + sizeof(int) * a.ld;
    if (n_stencils > 0x7fffffff) return fail(c, MF_ERR_ARGUMENT, "weights: too many stencils");
    hipLaunchKernelGGL(k_weights, dim3((unsigned)n_stencils), dim3(64), lds, c->stream, a);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// ---------------------------------------------------- operator packing
__global__ void k_pack_flux(int64_t n, const double* wx, const double* wy, double2* wxy, float2* w32)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    wxy[q] = make_double2(wx[q], wy[q]);
    w32[q] = make_float2((float)wx[q], (float)wy[q]);
}

// H = Dx^2k + Dy^2k, elementwise (hyperviscosity.jl:44)
__global__ void k_pack_hv(int64_t n, const double* wx, const double* wy, double* hv, float* h32)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double h = wy ? wx[q] + wy[q] : wx[q];
    hv[q] = h;
    h32[q] = (float)h;
}

int launch_pack_ops(mf_ctx* c, int target, const double* d_wx, const double* d_wy)
{
    const int64_t n = c->N * c->k;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (target == MF_OP_FLUX)
        hipLaunchKernelGGL(k_pack_flux, dim3(g), dim3(256), 0, c->stream, n, d_wx, d_wy, c->d_wxy, c->d_wxy32);
    else if (target == MF_OP_HYPERVISCOSITY)
        hipLaunchKernelGGL(k_pack_hv, dim3(g), dim3(256), 0, c->stream, n, d_wx, d_wy, c->d_hv, c->d_hv32);
    else
        hipLaunchKernelGGL(k_pack_hv, dim3(g), dim3(256), 0, c->stream, n, d_wx, d_wy, c->d_lap, c->d_lap32);
    MF_HIP(c, hipGetLastError());
    c->tm_dirty = true;
    return MF_OK;
}

// ------------------------------------------------- transposed operator
// Column i of D lists the rows r whose stencil contains i; stored as
// SELL-64 (slices of 64 consecutive columns, padded to the slice's
// largest count, entry e of column i at off[slice] + e*64 + i%64), each
// column sorted by the caller's row index like the reference's CSC
// (compute_operators.jl:443-452).
__global__ void k_t_count(int64_t N, int K, const int32_t* idx, int32_t* deg)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * K) return;
    atomicAdd(&deg[idx[q]], 1);
}

__global__ void k_t_width(int64_t N, int64_t ns, const int32_t* deg, int32_t* w, int64_t* wsize)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    int mx = 0;
    for (int l = 0; l < kSellC; ++l) {
        const int64_t i = s * kSellC + l;
        if (i < N) mx = max(mx, deg[i]);
    }
    w[s] = mx;
    wsize[s] = (int64_t)mx * kSellC;
}

__global__ void k_t_fill(int64_t N, int K, const int32_t* idx, const double2* wxy, const int64_t* off,
                         int32_t* cursor, int32_t* trow, double2* tw, int64_t* tsrc)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * K) return;
    const int64_t r = q % N;
    const int32_t j = idx[q];
    const int pos = atomicAdd(&cursor[j], 1);
    const int64_t t = off[j / kSellC] + (int64_t)pos * kSellC + (j % kSellC);
    trow[t] = (int32_t)r;
    tw[t] = wxy[q];
    tsrc[t] = q;
}

__global__ void k_t_sort(int64_t N, const int32_t* deg, const int32_t* w, const int64_t* off,
                         const int32_t* uid, int32_t* trow, double2* tw, float2* tw32, int64_t* tsrc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t base = off[i / kSellC] + (i % kSellC);
    const int n = deg[i];
    for (int a2 = 1; a2 < n; ++a2) {
        const int32_t rv = trow[base + (int64_t)a2 * kSellC];
        const int32_t rk = uid[rv];
        const double2 wv = tw[base + (int64_t)a2 * kSellC];
        const int64_t sv = tsrc[base + (int64_t)a2 * kSellC];
        int b = a2 - 1;
        while (b >= 0 && uid[trow[base + (int64_t)b * kSellC]] > rk) {
            trow[base + (int64_t)(b + 1) * kSellC] = trow[base + (int64_t)b * kSellC];
            tw[base + (int64_t)(b + 1) * kSellC] = tw[base + (int64_t)b * kSellC];
            tsrc[base + (int64_t)(b + 1) * kSellC] = tsrc[base + (int64_t)b * kSellC];
            --b;
        }
        trow[base + (int64_t)(b + 1) * kSellC] = rv;
        tw[base + (int64_t)(b + 1) * kSellC] = wv;
        tsrc[base + (int64_t)(b + 1) * kSellC] = sv;
    }
    const int W = w[i / kSellC];
    for (int a2 = n; a2 < W; ++a2) {
        trow[base + (int64_t)a2 * kSellC] = (int32_t)i;
        tw[base + (int64_t)a2 * kSellC] = make_double2(0.0, 0.0);
        tsrc[base + (int64_t)a2 * kSellC] = -1;
    }
    for (int a2 = 0; a2 < W; ++a2) {
        const double2 v = tw[base + (int64_t)a2 * kSellC];
        tw32[base + (int64_t)a2 * kSellC] = make_float2((float)v.x, (float)v.y);
    }
}

int launch_build_transpose(mf_ctx* c)
{
    const int64_t N = c->N, K = c->k;
    const int64_t ns = (N + kSellC - 1) / kSellC;
    int32_t *deg = nullptr, *cursor = nullptr;
    int64_t* wsize = nullptr;
    int rc;
    dfree(c->d_sell_off);
    dfree(c->d_sell_w);
    dfree(c->d_trow);
    dfree(c->d_tw);
    dfree(c->d_tw32);
    dfree(c->d_tsrc);
    dfree(c->d_tl);
    dfree(c->d_tl32);
    c->d_sell_off = nullptr;
    c->d_sell_w = nullptr;
    c->d_trow = nullptr;
    c->d_tw = nullptr;
    c->d_tw32 = nullptr;
    c->d_tsrc = nullptr;
    c->d_tl = nullptr;
    c->d_tl32 = nullptr;
    c->lap_T = false;
    if ((rc = dalloc(c, &deg, N)) || (rc = dalloc(c, &cursor, N)) || (rc = dalloc(c, &wsize, ns + 1)) ||
        (rc = dalloc(c, &c->d_sell_off, ns + 1)) || (rc = dalloc(c, &c->d_sell_w, ns)))
        return rc;
    MF_HIP(c, hipMemsetAsync(deg, 0, sizeof(int32_t) * N, c->stream));
    MF_HIP(c, hipMemsetAsync(cursor, 0, sizeof(int32_t) * N, c->stream));
    MF_HIP(c, hipMemsetAsync(wsize, 0, sizeof(int64_t) * (ns + 1), c->stream));
    const unsigned gq = (unsigned)((N * K + 255) / 256), gs = (unsigned)((ns + 255) / 256);
    hipLaunchKernelGGL(k_t_count, dim3(gq), dim3(256), 0, c->stream, N, (int)K, c->d_idx, deg);
    hipLaunchKernelGGL(k_t_width, dim3(gs), dim3(256), 0, c->stream, N, ns, deg, c->d_sell_w, wsize);
    size_t tb = 0;
    MF_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, wsize, c->d_sell_off, (int)(ns + 1), c->stream));
    void* tmp = nullptr;
    MF_HIP(c, hipMalloc(&tmp, tb + 16));
    MF_HIP(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb, wsize, c->d_sell_off, (int)(ns + 1), c->stream));
    int64_t total = 0;
    MF_HIP(c, hipMemcpyAsync(&total, c->d_sell_off + ns, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    (void)hipFree(tmp);
    if ((rc = dalloc(c, &c->d_trow, total)) || (rc = dalloc(c, &c->d_tw, total)) ||
        (rc = dalloc(c, &c->d_tw32, total)) || (rc = dalloc(c, &c->d_tsrc, total)))
        return rc;
    // entries of lanes past N in the last slice are never filled: mark
    // every entry as padding first
    MF_HIP(c, hipMemsetAsync(c->d_tsrc, 0xff, sizeof(int64_t) * total, c->stream));
    hipLaunchKernelGGL(k_t_fill, dim3(gq), dim3(256), 0, c->stream, N, (int)K, c->d_idx, c->d_wxy,
                       c->d_sell_off, cursor, c->d_trow, c->d_tw, c->d_tsrc);
    hipLaunchKernelGGL(k_t_sort, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, N, deg,
                       c->d_sell_w, c->d_sell_off, c->d_uid, c->d_trow, c->d_tw, c->d_tw32, c->d_tsrc);
    MF_HIP(c, hipGetLastError());
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(deg);
    dfree(cursor);
    dfree(wsize);
    c->n_slices = ns;
    c->t_entries = total;
    c->have_T = true;
    return MF_OK;
}

// int16 neighbour offsets j - i, when every one fits
__global__ void k_didx(int64_t N, int K, const int32_t* idx, int16_t* didx, int* bad)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * K) return;
    const int64_t i = q % N;
    const int64_t dlt = (int64_t)idx[q] - i;
    if (dlt < -32768 || dlt > 32767) {
        atomicOr(bad, 1);
        didx[q] = 0;
    } else {
        didx[q] = (int16_t)dlt;
    }
}

int launch_build_didx(mf_ctx* c)
{
    const int64_t n = c->N * c->k;
    int* bad = nullptr;
    int rc = dalloc(c, &bad, 1);
    if (rc) return rc;
    MF_HIP(c, hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_didx, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->N, c->k,
                       c->d_idx, c->d_didx, bad);
    int hb = 0;
    MF_HIP(c, hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    dfree(bad);
    c->didx_ok = (hb == 0);
    return MF_OK;
}

// ------------------------------------------ column-sorted stencil copy
// Sorted slot order of every node's stencil (ascending column index in
// the caller's numbering), the order in which the reference's CSC products
// accumulate a row.
__global__ void k_sort_stencil(int64_t N, int K, const int32_t* idx, const int32_t* uid, int32_t* idx_s,
                               uint8_t* perm)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int32_t key[kMaxStencilM];
    uint8_t pos[kMaxStencilM];
    for (int s = 0; s < K; ++s) {
        key[s] = uid[idx[(int64_t)s * N + i]];
        pos[s] = (uint8_t)s;
    }
    for (int a2 = 1; a2 < K; ++a2) {
        const int32_t kv = key[a2];
        const uint8_t pv = pos[a2];
        int b = a2 - 1;
        while (b >= 0 && key[b] > kv) {
            key[b + 1] = key[b];
            pos[b + 1] = pos[b];
            --b;
        }
        key[b + 1] = kv;
        pos[b + 1] = pv;
    }
    for (int s = 0; s < K; ++s) {
        idx_s[(int64_t)s * N + i] = idx[(int64_t)pos[s] * N + i];
        perm[(int64_t)s * N + i] = pos[s];
    }
}

__global__ void k_permute_slots(int64_t N, int K, const uint8_t* perm, const double* src, double* dst)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * K) return;
    const int64_t i = q % N;
    dst[q] = src[(int64_t)perm[q] * N + i];
}

__global__ void k_permute_sum(int64_t N, int K, const uint8_t* perm, const double* a, const double* b,
                              double* dst)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * K) return;
    const int64_t i = q % N;
    const int64_t src = (int64_t)perm[q] * N + i;
    dst[q] = b ? a[src] + b[src] : a[src];
}

int launch_sort_stencil(mf_ctx* c)
{
    const int64_t N = c->N;
    hipLaunchKernelGGL(k_sort_stencil, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, N, c->k,
                       c->d_idx, c->d_uid, c->d_idx_s, c->d_perm_s);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// Column-sorted copies of the operators: Dx, Dy separately; H = Dx^2k + Dy^2k.
int launch_pack_sorted(mf_ctx* c, int target, const double* d_wx, const double* d_wy)
{
    const int64_t n = c->N * c->k;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (target == MF_OP_FLUX) {
        hipLaunchKernelGGL(k_permute_slots, dim3(g), dim3(256), 0, c->stream, c->N, c->k, c->d_perm_s, d_wx,
                           c->d_wx_s);
        hipLaunchKernelGGL(k_permute_slots, dim3(g), dim3(256), 0, c->stream, c->N, c->k, c->d_perm_s, d_wy,
                           c->d_wy_s);
    } else {
        hipLaunchKernelGGL(k_permute_sum, dim3(g), dim3(256), 0, c->stream, c->N, c->k, c->d_perm_s, d_wx, d_wy,
                           target == MF_OP_HYPERVISCOSITY ? c->d_hv_s : c->d_lap_s);
    }
    MF_HIP(c, hipGetLastError());
    c->tm_dirty = true;
    return MF_OK;
}

// lap of the transposed entries (Tominec source): w[e] = lap at the
// slot-major stencil position src[e], 0 for padding (src[e] = -1).
__global__ void k_fill_lap_t(int64_t n, const int64_t* src, const double* lap, int64_t n_lap, double* w,
                             float* w32)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int64_t q = src[e];
    const double v = (q >= 0 && q < n_lap) ? lap[q] : 0.0;
    w[e] = v;
    w32[e] = (float)v;
}

int launch_fill_lap_t(mf_ctx* c, int64_t n, const int64_t* d_src, double* d_w, float* d_w32)
{
    if (n <= 0) return MF_OK;
    hipLaunchKernelGGL(k_fill_lap_t, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, d_src,
                       c->d_lap, c->N * (int64_t)c->k, d_w, d_w32);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// ------------------------------------------------- layout conversions
// Between the caller's row-major arrays (node u, entry s at u*K + s; field
// v of node u at v*N + u) and the internal layouts (slot-major ELL, node-
// major state), with the node order applied; MAP also maps index values.
template <class T, bool MAP>
__global__ void k_ell_to_rows(int64_t N, int K, const T* ell, T* rows, const int32_t* inv, const int32_t* uid)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * K) return;
    const int64_t u = q / K, s = q % K;
    const T v = ell[s * N + inv[u]];
    if constexpr (MAP)
        rows[q] = uid[v];
    else
        rows[q] = v;
}

template <class T, bool MAP>
__global__ void k_rows_to_ell(int64_t N, int K, const T* rows, T* ell, const int32_t* uid, const int32_t* inv)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * K) return;
    const int64_t s = q / N, p = q % N;
    const T v = rows[(int64_t)uid[p] * K + s];
    if constexpr (MAP)
        ell[q] = inv[v];
    else
        ell[q] = v;
}

int launch_ell_to_rows(mf_ctx* c, const int32_t* d_ell, int32_t* d_rows, int k)
{
    const int64_t n = c->N * k;
    hipLaunchKernelGGL((k_ell_to_rows<int32_t, true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, c->N, k, d_ell, d_rows, c->d_inv, c->d_uid);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}
int launch_ell_to_rows_d(mf_ctx* c, const double* d_ell, double* d_rows, int k)
{
    const int64_t n = c->N * k;
    hipLaunchKernelGGL((k_ell_to_rows<double, false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, c->N, k, d_ell, d_rows, c->d_inv, c->d_uid);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}
int launch_rows_to_ell(mf_ctx* c, const int32_t* d_rows, int32_t* d_ell, int k)
{
    const int64_t n = c->N * k;
    hipLaunchKernelGGL((k_rows_to_ell<int32_t, true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, c->N, k, d_rows, d_ell, c->d_uid, c->d_inv);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}
int launch_rows_to_ell_d(mf_ctx* c, const double* d_rows, double* d_ell, int k)
{
    const int64_t n = c->N * k;
    hipLaunchKernelGGL((k_rows_to_ell<double, false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, c->N, k, d_rows, d_ell, c->d_uid, c->d_inv);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

__global__ void k_soa_to_aos(int64_t N, int V, const double* soa, double* aos, const int32_t* uid)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * V) return;
    const int64_t p = q / V, v = q % V;
    aos[q] = soa[v * N + uid[p]];
}

__global__ void k_aos_to_soa(int64_t N, int V, const double* aos, double* soa, const int32_t* inv)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * V) return;
    const int64_t v = q / N, u = q % N;
    soa[q] = aos[(int64_t)inv[u] * V + v];
}

int launch_soa_to_aos(mf_ctx* c, const double* d_soa, double* d_aos)
{
    const int64_t n = c->N * c->nvar;
    hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->N,
                       c->nvar, d_soa, d_aos, c->d_uid);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}
int launch_aos_to_soa(mf_ctx* c, const double* d_aos, double* d_soa)
{
    const int64_t n = c->N * c->nvar;
    hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->N,
                       c->nvar, d_aos, d_soa, c->d_inv);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// plain streaming copy of n doubles (n a multiple of 4 or not)
__global__ void k_copy_buffer(int64_t n, const double* __restrict__ src, double* __restrict__ dst)
{
    const int64_t n4 = n / 4;
    const double4* s4 = reinterpret_cast<const double4*>(src);
    double4* d4 = reinterpret_cast<double4*>(dst);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) d4[q] = s4[q];
    for (int64_t q = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = src[q];
}

int launch_copy_buffer(mf_ctx* c, const double* d_src, double* d_dst, int64_t n)
{
    const int64_t blocks = std::min<int64_t>(4096, (n / 4 + 255) / 256 + 1);
    hipLaunchKernelGGL(k_copy_buffer, dim3((unsigned)blocks), dim3(256), 0, c->stream, n, d_src, d_dst);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// dst[i] = src[(i & ~255) | reverse8(i & 255)]: every element of src read
// once and every element of dst written once (16 bytes per element, like the
// copy), but each wavefront's loads are scattered over a 2 KiB window, the
// access shape of a stencil gather.  A second known-bytes kernel for the
// HBM traffic counters.
__global__ void k_gather_probe(int64_t n, const double* __restrict__ src, double* __restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t full = n & ~(int64_t)255;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const int64_t r = q < full ? (q & ~(int64_t)255) | (int64_t)(__builtin_bitreverse32((uint32_t)(q & 255)) >> 24)
                                   : q;
        dst[q] = src[r];
    }
}

int launch_gather_probe(mf_ctx* c, const double* d_src, double* d_dst, int64_t n)
{
    const int64_t blocks = std::min<int64_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_gather_probe, dim3((unsigned)blocks), dim3(256), 0, c->stream, n, d_src, d_dst);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// read-modify-write of a buffer larger than the last-level cache, so that a
// call timed after it finds none of its own data cached
__global__ void k_flush(int64_t n, double* p)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) p[q] = p[q] * 0.5 + 1.0;
}

int launch_flush(mf_ctx* c, double* d_buf, int64_t n)
{
    hipLaunchKernelGGL(k_flush, dim3(4096), dim3(256), 0, c->stream, n, d_buf);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

// per-node arrays back to the caller's order: dst[u] = src[inv[u]]
template <class T>
__global__ void k_gather_nodes(int64_t N, const int32_t* inv, const T* src, T* dst)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < N) dst[u] = src[inv[u]];
}

int launch_gather_nodes_d(mf_ctx* c, const double* d_src, double* d_dst)
{
    hipLaunchKernelGGL(k_gather_nodes<double>, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream,
                       c->N, c->d_inv, d_src, d_dst);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}
int launch_gather_nodes_i8(mf_ctx* c, const int8_t* d_src, int8_t* d_dst)
{
    hipLaunchKernelGGL(k_gather_nodes<int8_t>, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream,
                       c->N, c->d_inv, d_src, d_dst);
    MF_HIP(c, hipGetLastError());
    return MF_OK;
}

}  // namespace mf
