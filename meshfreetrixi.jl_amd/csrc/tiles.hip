// LDS tiles of the stencil operators (built once per stencil / operator).
//
// The neighbour gathers of the rhs! kernels touch one cache line per lane
// per load: a wavefront's 64 lanes read 64 scattered 32- or 64-byte node
// records.  With the nodes in Morton order, the 256 nodes of a workgroup
// form a compact patch whose stencils reach only ~2x as many distinct
// nodes.  Each workgroup therefore loads that halo once into LDS (runs of
// consecutive node records) and the stencil loops read LDS through 16-bit
// halo positions stored in place of the 32-bit node indices.
//
// Phase A (flux divergence, sources, viscosity phase A): halo = the union
// of the tile's stencils; positions, and copies of the operators, in the
// tile-major layout (tm_pos: a tile's K slots back to back), so that one
// tile's entries of an array share a few pages: with the slot-major layout
// (slot stride N) a tile touched K pages per array, and on multi-million
// node sets the address translation (UTCL2 busy 91% of the time) limited
// the stencil kernel.
// Phase B (transposed gathers D^T G, lap^T L): rows of the tile sorted by their
// transposed degree (long rows first) so that the 64 rows of a wavefront
// have similar lengths, stored SELL-64 per wavefront; entries of a row in
// ascending caller index (the reference's CSC order); halo = the union of
// the rows' entries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <type_traits>
#include <vector>

#include "mf_device.h"

namespace mf {

constexpr int kTile = 256;
constexpr int kSlicesPerTile = kTile / kSellC;  // 4
constexpr size_t kTileLdsBytes = 48 * 1024;     // largest phase-A halo staged in LDS
constexpr size_t kTileLdsBytesB = 64 * 1024;    // phase B (records of G and lap u)

void free_tiles_b(mf_ctx* c)
{
    dfree(c->d_hb_off);
    dfree(c->d_hb_ids);
    dfree(c->d_sb_off);
    dfree(c->d_sb_w);
    dfree(c->d_rowb);
    dfree(c->d_lb);
    dfree(c->d_twb);
    dfree(c->d_twb32);
    dfree(c->d_tb_src);
    dfree(c->d_twl);
    dfree(c->d_twl32);
    c->d_tb_src = nullptr;
    c->d_twl = nullptr;
    c->d_twl32 = nullptr;
    c->lap_tb = false;
    c->tiles_b_rec = 0;
    c->d_hb_off = c->d_sb_off = nullptr;
    c->d_hb_ids = c->d_sb_w = nullptr;
    c->d_lb = nullptr;
    c->d_rowb = nullptr;
    c->d_twb = nullptr;
    c->d_twb32 = nullptr;
    c->tiles_b = false;
    c->tried_b = false;
    c->hb_max = 0;
}

static void free_tile_ops(mf_ctx* c)
{
    dfree(c->d_wxy_t);
    dfree(c->d_wxy32_t);
    dfree(c->d_hv_t);
    dfree(c->d_hv32_t);
    dfree(c->d_lap_t);
    dfree(c->d_lap32_t);
    dfree(c->d_wx_s_t);
    dfree(c->d_wy_s_t);
    dfree(c->d_hv_s_t);
    dfree(c->d_lap_s_t);
    c->d_wxy_t = nullptr;
    c->d_wxy32_t = nullptr;
    c->d_hv_t = c->d_lap_t = c->d_wx_s_t = c->d_wy_s_t = c->d_hv_s_t = c->d_lap_s_t = nullptr;
    c->d_hv32_t = c->d_lap32_t = nullptr;
    c->tm_dirty = true;
}

void free_tiles(mf_ctx* c)
{
    free_tiles_b(c);
    free_tile_ops(c);
    dfree(c->d_ha_off);
    dfree(c->d_ha_ids);
    dfree(c->d_la);
    dfree(c->d_la_s);
    c->d_ha_off = nullptr;
    c->d_ha_ids = nullptr;
    c->d_la = nullptr;
    c->d_la_s = nullptr;
    c->tiles_a = false;
    c->tried_a = false;
    c->n_tiles = 0;
    c->ha_max = 0;
}

// (Dx, Dy) of each transposed entry from the forward operator: src[e] is
// the slot-major position s*N + j of the entry, or -1 for padding.
__global__ void k_fill_tile_weights(int64_t n, const int64_t* src, const double2* wxy, double2* tw, float2* tw32)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int64_t q = src[e];
    const double2 w = q >= 0 ? wxy[q] : make_double2(0.0, 0.0);
    tw[e] = w;
    tw32[e] = make_float2((float)w.x, (float)w.y);
}

// Slot-major (s N + i) operator array to the tile-major layout (tm_pos);
// the padding entries of the last tile are zero.
template <class T>
__global__ void k_to_tile_major(int64_t N, int K, int64_t n_out, const T* src, T* dst)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_out) return;
    const int64_t r = q % kTile, ts = q / kTile;  // ts = t K + s
    const int64_t t = ts / K, sl = ts % K;
    const int64_t i = t * kTile + r;
    dst[q] = i < N ? src[sl * N + i] : T{};
}

// The tile-major copies of every operator array present, after the tiles
// are built and whenever an operator changed since the last copy.
int ensure_tile_ops(mf_ctx* c)
{
    if (!c->tiles_a || !c->tm_dirty) return MF_OK;
    const int64_t n_out = c->n_tiles * (int64_t)c->k * kTile;
    int rc;
    auto copy = [&](auto* src, auto** dst) -> int {
        using T = std::remove_pointer_t<decltype(src)>;
        if (!src) return MF_OK;
        if (!*dst && (rc = dalloc(c, dst, (size_t)n_out))) return rc;
        hipLaunchKernelGGL(k_to_tile_major<T>, dim3((unsigned)((n_out + 255) / 256)), dim3(256), 0, c->stream,
                           c->N, c->k, n_out, (const T*)src, *dst);
        MF_HIP(c, hipGetLastError());
        return MF_OK;
    };
    if ((rc = copy(c->d_wxy, &c->d_wxy_t)) || (rc = copy(c->d_wxy32, &c->d_wxy32_t)) ||
        (rc = copy(c->d_hv, &c->d_hv_t)) || (rc = copy(c->d_hv32, &c->d_hv32_t)) ||
        (rc = copy(c->d_lap, &c->d_lap_t)) || (rc = copy(c->d_lap32, &c->d_lap32_t)) ||
        (rc = copy(c->d_wx_s, &c->d_wx_s_t)) || (rc = copy(c->d_wy_s, &c->d_wy_s_t)) ||
        (rc = copy(c->d_hv_s, &c->d_hv_s_t)) || (rc = copy(c->d_lap_s, &c->d_lap_s_t)))
        return rc;
    c->tm_dirty = false;
    return MF_OK;
}

template <class T>
static int upload(mf_ctx* c, T** dst, const std::vector<T>& v)
{
    int rc = dalloc(c, dst, v.size());
    if (rc) return rc;
    if (!v.empty())
        MF_HIP(c, hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream));
    return MF_OK;
}

// rec_b: doubles per phase-B halo record (G: 2 nvar, L: nvar), which
// decides whether a tile's phase-B halo fits in LDS.
int build_tiles(mf_ctx* c, bool want_b, int rec_b)
{
    const int64_t N = c->N;
    const int K = c->k;
    const int V = c->nvar;
    const int64_t nt = (N + kTile - 1) / kTile;
    int rc;
    std::vector<int32_t> idx((size_t)N * K), idx_s((size_t)N * K);
    MF_HIP(c, hipMemcpyAsync(idx.data(), c->d_idx, sizeof(int32_t) * idx.size(), hipMemcpyDeviceToHost,
                             c->stream));
    MF_HIP(c, hipMemcpyAsync(idx_s.data(), c->d_idx_s, sizeof(int32_t) * idx_s.size(), hipMemcpyDeviceToHost,
                             c->stream));
    MF_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> mark(N, -1);
    std::vector<int32_t> ids;
    ids.reserve(4 * kTile);

    if (!c->tried_a) {
        free_tiles(c);
        c->tried_a = true;
        std::vector<int64_t> off(nt + 1, 0);
        std::vector<int32_t> all;
        all.reserve((size_t)N * 2);
        // tile-major (tm_pos): tile t's slot s at (t K + s) kTile
        std::vector<uint16_t> la((size_t)nt * K * kTile, 0), la_s((size_t)nt * K * kTile, 0);
        int hmax = 0;
        bool fits = true;
        for (int64_t t = 0; t < nt && fits; ++t) {
            const int64_t i0 = t * kTile, i1 = std::min<int64_t>(N, i0 + kTile);
            ids.clear();
            for (int s = 0; s < K; ++s)
                for (int64_t i = i0; i < i1; ++i) {
                    const int32_t j = idx[(size_t)s * N + i];
                    if (mark[j] < 0) {
                        mark[j] = 0;
                        ids.push_back(j);
                    }
                }
            std::sort(ids.begin(), ids.end());
            // phase A halo records: 6 doubles per node for Euler (state, v1, v2)
            const size_t rec = (V == 4) ? 6 : (size_t)V;
            if (ids.size() > 65535 || ids.size() * rec * sizeof(double) > kTileLdsBytes) fits = false;
            for (size_t q = 0; q < ids.size(); ++q) mark[ids[q]] = (int32_t)q;
            for (int s = 0; s < K; ++s)
                for (int64_t i = i0; i < i1; ++i) {
                    const size_t tq = (size_t)(t * K + s) * kTile + (size_t)(i - i0);
                    la[tq] = (uint16_t)mark[idx[(size_t)s * N + i]];
                    la_s[tq] = (uint16_t)mark[idx_s[(size_t)s * N + i]];
                }
            for (int32_t j : ids) mark[j] = -1;
            hmax = std::max<int>(hmax, (int)ids.size());
            all.insert(all.end(), ids.begin(), ids.end());
            off[t + 1] = (int64_t)all.size();
        }
        if (!fits) return MF_OK;  // global-memory kernels
        if ((rc = upload(c, &c->d_ha_off, off)) || (rc = upload(c, &c->d_ha_ids, all)) ||
            (rc = upload(c, &c->d_la, la)) || (rc = upload(c, &c->d_la_s, la_s)))
            return rc;
        MF_HIP(c, hipStreamSynchronize(c->stream));
        c->n_tiles = nt;
        c->ha_max = hmax;
        c->ha_total = (int64_t)all.size();
        c->tiles_a = true;
    }
    if (!want_b || !c->tiles_a || c->tried_b) return MF_OK;
    free_tiles_b(c);
    c->tried_b = true;
    c->tiles_b_rec = rec_b;

    // transposed adjacency: the entries (j, s) with idx[s][j] = i, per row i,
    // in ascending caller index of j
    std::vector<int32_t> uid(N);
    for (int64_t u = 0; u < N; ++u) uid[c->h_inv[u]] = (int32_t)u;
    std::vector<int64_t> rp(N + 1, 0);
    for (size_t q = 0; q < idx.size(); ++q) ++rp[idx[q] + 1];
    for (int64_t i = 0; i < N; ++i) rp[i + 1] += rp[i];
    std::vector<int64_t> ent(idx.size());  // slot-major position s*N + j
    {
        std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
        for (int s = 0; s < K; ++s)
            for (int64_t j = 0; j < N; ++j) {
                const int32_t i = idx[(size_t)s * N + j];
                ent[cur[i]++] = (int64_t)s * N + j;
            }
    }
    for (int64_t i = 0; i < N; ++i)
        std::sort(ent.begin() + rp[i], ent.begin() + rp[i + 1], [&](int64_t a, int64_t b) {
            const int32_t ua = uid[a % N], ub = uid[b % N];
            return ua < ub || (ua == ub && a < b);
        });
    std::vector<int64_t> hoff(nt + 1, 0), soff(nt * kSlicesPerTile + 1, 0);
    std::vector<int32_t> hall, sw(nt * kSlicesPerTile, 0);
    std::vector<uint8_t> rowb((size_t)nt * kTile, 0);
    std::vector<uint16_t> lb;
    std::vector<int64_t> src;
    int hmax = 0;
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t i0 = t * kTile, i1 = std::min<int64_t>(N, i0 + kTile);
        ids.clear();
        for (int64_t i = i0; i < i1; ++i)
            for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
                const int32_t j = (int32_t)(ent[e] % N);
                if (mark[j] < 0) {
                    mark[j] = 0;
                    ids.push_back(j);
                }
            }
        std::sort(ids.begin(), ids.end());
        if (ids.size() > 65535 || ids.size() * (size_t)rec_b * sizeof(double) > kTileLdsBytesB) {
            for (int32_t j : ids) mark[j] = -1;
            return MF_OK;  // global-memory transposed gather
        }
        for (size_t q = 0; q < ids.size(); ++q) mark[ids[q]] = (int32_t)q;
        // rows by degree, longest first (rows past N have degree 0)
        int order[kTile];
        int deg[kTile];
        for (int r = 0; r < kTile; ++r) {
            order[r] = r;
            deg[r] = (i0 + r < i1) ? (int)(rp[i0 + r + 1] - rp[i0 + r]) : 0;
        }
        std::stable_sort(order, order + kTile, [&](int a, int b) { return deg[a] > deg[b]; });
        for (int p = 0; p < kTile; ++p) rowb[(size_t)t * kTile + p] = (uint8_t)order[p];
        for (int w = 0; w < kSlicesPerTile; ++w) {
            int W = 0;
            for (int l = 0; l < kSellC; ++l) W = std::max(W, deg[order[w * kSellC + l]]);
            const int64_t sl = t * kSlicesPerTile + w;
            sw[sl] = W;
            const size_t base = lb.size();
            lb.resize(base + (size_t)W * kSellC, 0);
            src.resize(base + (size_t)W * kSellC, -1);
            for (int l = 0; l < kSellC; ++l) {
                const int r = order[w * kSellC + l];
                if (i0 + r >= i1) continue;
                const int64_t i = i0 + r;
                for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
                    const size_t at = base + (size_t)(e - rp[i]) * kSellC + l;
                    lb[at] = (uint16_t)mark[ent[e] % N];
                    src[at] = ent[e];
                }
            }
            soff[sl + 1] = (int64_t)lb.size();
        }
        for (int32_t j : ids) mark[j] = -1;
        hmax = std::max<int>(hmax, (int)ids.size());
        hall.insert(hall.end(), ids.begin(), ids.end());
        hoff[t + 1] = (int64_t)hall.size();
    }
    if ((rc = upload(c, &c->d_hb_off, hoff)) || (rc = upload(c, &c->d_hb_ids, hall)) ||
        (rc = upload(c, &c->d_sb_off, soff)) || (rc = upload(c, &c->d_sb_w, sw)) ||
        (rc = upload(c, &c->d_rowb, rowb)) || (rc = upload(c, &c->d_lb, lb)) ||
        (rc = upload(c, &c->d_tb_src, src)) || (rc = dalloc(c, &c->d_twb, lb.size())) ||
        (rc = dalloc(c, &c->d_twb32, lb.size())))
        return rc;
    const int64_t ne = (int64_t)lb.size();
    hipLaunchKernelGGL(k_fill_tile_weights, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream, ne,
                       c->d_tb_src, c->d_wxy, c->d_twb, c->d_twb32);
    MF_HIP(c, hipGetLastError());
    MF_HIP(c, hipStreamSynchronize(c->stream));
    c->hb_max = hmax;
    c->hb_total = (int64_t)hall.size();
    c->nb_entries = ne;
    c->tiles_b = true;
    return MF_OK;
}

}  // namespace mf
