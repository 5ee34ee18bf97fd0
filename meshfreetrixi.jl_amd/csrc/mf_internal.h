// Internal declarations shared by the libmfrbf sources (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mfrbf.h"

namespace mf {

constexpr int kMaxVar = 4;
constexpr int kMaxTags = 4096;
constexpr int kMaxHist = 8;
constexpr int kSellC = 64;  // slice height of the transposed operator
constexpr int kMaxStencilM = 64;  // saddle systems up to 64 x 64 (one wavefront)

// Per-step scalars kept in device memory so that captured graphs read the
// current time, history weights and success_iter at replay.
struct StepScalars {
    double t;                    // integrator time t_n
    double time_hist[kMaxHist];  // most recent first
    double tw[kMaxHist];         // time-derivative weights
    double norms[2 * kMaxVar];   // mean(u), ninf of |u - mean| (residual viscosity)
    int success_iter;
    int hist_head;  // ring slot of the most recent snapshot
    int q;          // number of history points in use
    int hist_si;  // success_iter as last seen by the history callback
    int pad[3];
};

// Conjugate-gradient scalars of the IGR source, in device memory.
struct IgrScalars {
    double res[2];  // residual norm of CG iteration it in slot it & 1
    double tol;     // sqrt(eps) |b|
    int done;       // the current iteration is past convergence
    int pad;
};

// Term table of the k-th derivative of r^p along one coordinate s:
// sum_t c_t s^a_t r^b_t.
struct PhsTable {
    int nt;
    int a[16];
    int b[16];
    double c[16];
};
PhsTable phs_table(int p, int k);

struct BcDev {
    int ntags;
    int nb;
    int any_slip;
    int multi;  // some node belongs to more than one tag
    int32_t* bidx = nullptr;     // nb
    int32_t* btag = nullptr;     // nb
    double* bnormal = nullptr;   // 2 nb
    double* btable = nullptr;    // nvar nb
    uint8_t* flag = nullptr;     // N: bit0 Dirichlet, bit1 slip
    int32_t* entry = nullptr;    // N: boundary entry of the node, -1 if none
    int32_t* ecode = nullptr;    // nb: first-pass kind | function << 4 | tag << 8 of each entry
    std::vector<int32_t> tag_kind;
    std::vector<int32_t> tag_fn;
    std::vector<int32_t> tag_ptr;
    std::vector<double> tag_params;    // 8 per tag
    int32_t* d_kind = nullptr;         // ntags: tag kinds
    int32_t* d_kind_slip = nullptr;    // ntags: slip walls kept, other kinds as do-nothing
    int32_t* d_fn = nullptr;           // ntags: value functions
    double* d_params = nullptr;        // 8 ntags: value-function parameters
};

}  // namespace mf

struct mf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int cu_count = 0;

    // cloud, in the internal node order: internal node p is the caller's
    // node d_uid[p]; d_inv is the inverse map.  Every array below is
    // indexed by internal node; the ABI converts at its edges.
    int64_t N = 0;
    double* d_x = nullptr;
    double* d_y = nullptr;
    int32_t* d_uid = nullptr;
    int32_t* d_inv = nullptr;
    std::vector<int32_t> h_inv;  // host copy of d_inv (boundary and stencil indices)
    int node_order = 0;          // MF_TUNE_NODE_ORDER value for the next mf_set_points
    int n_groups = 1;            // independent clouds in the point set (mf_set_points_grouped)
    std::vector<int64_t> group_ptr;  // n_groups + 1: each cloud's node range (caller's numbering)

    // stencil (ELL, slot-major: entry s of node i at s*N + i)
    int k = 0;
    int32_t* d_idx = nullptr;
    int16_t* d_didx = nullptr;  // idx - i as int16 (split precision), if it fits
    bool didx_ok = false;
    // the same stencil with each node's entries sorted by column (the
    // reference's CSC order) and the slot permutation that sorts it
    int32_t* d_idx_s = nullptr;
    uint8_t* d_perm_s = nullptr;
    double* d_wx_s = nullptr;
    double* d_wy_s = nullptr;
    double* d_hv_s = nullptr;
    double* d_lap_s = nullptr;

    // operators
    double2* d_wxy = nullptr;  // (Dx, Dy)
    double* d_hv = nullptr;    // Dx^2k + Dy^2k
    float2* d_wxy32 = nullptr;
    float* d_hv32 = nullptr;
    double* d_lap = nullptr;  // Dxx + Dyy (Tominec source)
    float* d_lap32 = nullptr;
    bool have_flux = false, have_hv = false, have_lap = false;
    // transposed flux operator, SELL-64
    int64_t n_slices = 0;
    int64_t t_entries = 0;  // entries with padding
    int64_t* d_sell_off = nullptr;  // n_slices + 1
    int32_t* d_sell_w = nullptr;    // n_slices
    int32_t* d_trow = nullptr;
    double2* d_tw = nullptr;
    float2* d_tw32 = nullptr;
    int64_t* d_tsrc = nullptr;  // slot-major stencil position of each entry (-1: padding)
    double* d_tl = nullptr;     // lap of the transposed entries (Tominec)
    float* d_tl32 = nullptr;
    bool have_T = false;
    bool lap_T = false;  // d_tl filled from the current lap and transpose

    // physics
    int eq = MF_EQ_ADVECTION_2D;
    int nvar = 1;
    double p0 = 1.0, p1 = 1.0;  // advection velocity or (gamma, -)
    mf::BcDev bc;
    bool have_bc = false;
    double hv_gamma = 0.0;
    int visc_mode = MF_VISC_NONE;
    double c_rv = 1.0, c_uw = 1.0, h_visc = 0.0;
    int approx_order = 0;
    int hv_first = 1;
    double tom_gamma = 0.0;
    double igr_alpha = 0.0;  // IGR source (0: none)
    int igr_maxiter = 20;
    // rank of each source kind (MF_SRC_FLYER, _TOMINEC, _VISCOSITY) in the
    // SourceTerms keyword order; smaller adds to du first
    int src_rank[5] = {0, 0, 1, 2, 3};

    // state
    double* d_buf[4] = {nullptr, nullptr, nullptr, nullptr};
    double* d_S = nullptr;
    double* d_T = nullptr;
    double* d_G = nullptr;
    float* d_du32 = nullptr;  // split-FP32 partial du between the two viscosity phases
    double* d_L = nullptr;    // lap u between the phases (Tominec); FP32 in split-FP32
    double* d_igr = nullptr;  // IGR: 8 fields of N (1/rho, b, sigma, r, p, c, t1, t2)
    double* d_igr_part = nullptr;          // per-block partial sums: r.r, then p.c
    mf::IgrScalars* d_igr_sc = nullptr;  // CG scalars on the device
    double* d_eps = nullptr;
    int8_t* d_epsc = nullptr;
    double* d_approx = nullptr;
    double* d_hist = nullptr;
    double* d_partials = nullptr;  // 4 n_psum sums, then 4 n_pmax lexicographic maxima
    double* d_stage_psum = nullptr;  // 4 sums per block of kBlock nodes: a stage's output (stepping)
    double* d_flush = nullptr;     // kFlushDoubles, allocated by the first mf_time_rhs_cold
    int n_psum = 0, n_pmax = 0;
    mf::StepScalars* d_sc = nullptr;
    int host_success_iter = 0;
    bool fsal_valid = false;
    bool fsal_sums = false;   // the stored FSAL stage carries its next boundary pass and sums
    bool graph_sums = false;  // the captured step passes them between stages

    // stage limiter
    int lim_n = 0;
    double lim_thr[2] = {0, 0};
    int lim_var[2] = {0, 0};
    double* d_lim = nullptr;  // limiter scratch: nvar N values, then N flag bytes

    // graph of one SSPRK33 step
    // graphs of one step and of kGraphSteps consecutive steps (fewer launches)
    hipGraphExec_t step_exec = nullptr;
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t multi_exec = nullptr;
    hipGraph_t multi_graph = nullptr;
    double graph_dt = 0.0;  // dt the graphs were captured with
    double fsal_dt = 0.0;   // dt of the last mf_ssprk33 call (its FSAL stage)
    int graph_prec = -1, graph_flags = -1;

    // ragged weights batch
    int64_t rn = 0, rnnz = 0;
    int rmaxn = 0;
    int rminn = 0;
    int64_t* d_rptr = nullptr;
    int32_t* d_ridx = nullptr;
    int32_t* d_rorder = nullptr;          // stencils grouped by size
    std::vector<int64_t> rclass_start;    // per size class: first entry of d_rorder
    std::vector<int> rclass_n;            // per size class: stencil size
    double* d_rwx = nullptr;
    double* d_rwy = nullptr;
    int* d_status = nullptr;

    // adjoint scratch
    double* d_adj = nullptr;
    int64_t adj_len = 0;

    // LDS tiles: tile t = internal nodes [256 t, 256 t + 256).  Phase A
    // (stencil kernel) stages the states of the tile's stencil halo in LDS
    // and addresses them by 16-bit halo positions; phase B (transposed
    // gather) stages G of its halo and walks rows sorted by degree.
    int use_tiles = 1;  // MF_TUNE_TILES
    bool tiles_a = false, tiles_b = false;  // built and fitting in LDS
    bool tried_a = false, tried_b = false;  // build attempted for this stencil / operator
    int64_t n_tiles = 0;
    int ha_max = 0, hb_max = 0;          // largest halo of a tile
    int64_t ha_total = 0, hb_total = 0;  // all tiles' halo entries
    int64_t nb_entries = 0;              // phase-B entries with padding
    int64_t* d_ha_off = nullptr;         // n_tiles + 1
    int32_t* d_ha_ids = nullptr;         // halo node ids, sorted per tile
    uint16_t* d_la = nullptr;            // halo positions of the d_idx entries, tile-major
    uint16_t* d_la_s = nullptr;          // the same for the column-sorted stencil d_idx_s
    // tile-major copies of the operators for the tile kernels (rebuilt from
    // the slot-major arrays after every operator change: tm_dirty)
    double2* d_wxy_t = nullptr;
    float2* d_wxy32_t = nullptr;
    double* d_hv_t = nullptr;
    float* d_hv32_t = nullptr;
    double* d_lap_t = nullptr;
    float* d_lap32_t = nullptr;
    double* d_wx_s_t = nullptr;
    double* d_wy_s_t = nullptr;
    double* d_hv_s_t = nullptr;
    double* d_lap_s_t = nullptr;
    bool tm_dirty = true;
    int64_t* d_hb_off = nullptr;
    int32_t* d_hb_ids = nullptr;
    int64_t* d_sb_off = nullptr;         // 4 slices (64 rows) per tile
    int32_t* d_sb_w = nullptr;
    uint8_t* d_rowb = nullptr;           // tile position -> tile row
    uint16_t* d_lb = nullptr;            // halo positions, slice layout
    double2* d_twb = nullptr;            // (Dx, Dy) of the transposed entries
    float2* d_twb32 = nullptr;
    int64_t* d_tb_src = nullptr;         // slot-major stencil position of each entry (-1: padding)
    double* d_twl = nullptr;             // lap of the transposed entries (Tominec)
    float* d_twl32 = nullptr;
    int tiles_b_rec = 0;                 // doubles per phase-B halo record the tiles were sized for
    bool lap_tb = false;                 // d_twl filled from the current lap

    // launch tuning (mf_set_tuning)
    // per precision (MF_PREC_*); 0 = the default
    int tune_flux_batch[3] = {0, 0, 0};       // stencil entries per load batch
    int tune_transpose_batch[3] = {0, 0, 0};  // transposed entries per load batch
    int weights_solver = 0;                   // MF_TUNE_WEIGHTS_SOLVER

    // when set, the rhs! kernels are dispatched with start/stop events:
    // probe[2 s], probe[2 s + 1] for slot s (see launch_k in rhs.hip)
    hipEvent_t* probe = nullptr;
    unsigned probe_used = 0;  // slots dispatched with events during the probed call
};

namespace mf {

constexpr int kGraphSteps = 16;  // SSPRK33 steps per multi-step graph

// error helpers
int fail(mf_ctx* c, int code, const std::string& msg);
int hip_fail(mf_ctx* c, hipError_t e, const char* what);
#define MF_HIP(c, expr)                                   \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return mf::hip_fail((c), _e, #expr); \
    } while (0)

template <class T>
int dalloc(mf_ctx* c, T** p, size_t count);
void dfree(void* p);

// XCD-aware launch helper: blocks are dealt round-robin to the 8 XCDs; the
// remap gives every XCD one contiguous range of logical blocks.
inline unsigned xcd_grid(int64_t nblocks) { return (unsigned)(((nblocks + 7) / 8) * 8); }

// kernels / launchers (defined in the .hip files)
int launch_knn(mf_ctx* c, int k, double* d_dist2);
// the same search with its intermediates in a static device array (no
// allocation): row per caller node, caller indices, distances
int launch_knn_noalloc(mf_ctx* c, int k, int32_t* d_idx, double* d_dist);
int64_t knn_noalloc_max_nodes();
// n_stencils stencils: all of them (d_order null) or d_order[base + b],
// b < n_stencils, each with at most maxn nodes (sizes the LDS)
int launch_weights(mf_ctx* c, const int64_t* d_ptr, const int32_t* d_idx, int ell_k,
                   int64_t n_stencils, int maxn, int p, int deg, int kd, double* d_wx,
                   double* d_wy, int* d_status, const int32_t* d_order = nullptr, int64_t base = 0);
// the same weights by the null-space form on the matrix cores: one
// workgroup per stencil, stencil d_sid[b] (or b when d_sid is null) for
// workgroup b of n_launch; maxn, the largest of their sizes, picks the
// variant (up to 32, 48 or 64 nodes); uniform_n is the size of every
// stencil of the launch when they all have the same one (else 0), which
// selects the kernels compiled for that shape
int launch_weights_ns(mf_ctx* c, const int64_t* d_ptr, const int32_t* d_idx, int ell_k, const int32_t* d_sid,
                      int64_t n_launch, int maxn, int uniform_n, int p, int deg, int kd, double* d_wx, double* d_wy,
                      int* d_status);
int launch_pack_ops(mf_ctx* c, int target, const double* d_wx, const double* d_wy);
int launch_build_transpose(mf_ctx* c);
int launch_dx_stats(mf_ctx* c, const double* d_dist2, double* dx_min, double* dx_avg);
int launch_ell_to_rows(mf_ctx* c, const int32_t* d_ell, int32_t* d_rows, int k);
int launch_ell_to_rows_d(mf_ctx* c, const double* d_ell, double* d_rows, int k);
int launch_rows_to_ell(mf_ctx* c, const int32_t* d_rows, int32_t* d_ell, int k);
int launch_rows_to_ell_d(mf_ctx* c, const double* d_rows, double* d_ell, int k);
int launch_soa_to_aos(mf_ctx* c, const double* d_soa, double* d_aos);
int launch_aos_to_soa(mf_ctx* c, const double* d_aos, double* d_soa);
int launch_build_didx(mf_ctx* c);
// d_group (caller order, may be null): group of each point; the internal
// order keeps each group contiguous (Morton order within the group)
int launch_node_order(mf_ctx* c, const double* d_xu, const double* d_yu, const int32_t* d_group = nullptr);
void free_tiles(mf_ctx* c);
void free_tiles_b(mf_ctx* c);
int launch_gather_nodes_d(mf_ctx* c, const double* d_src, double* d_dst);
int launch_copy_buffer(mf_ctx* c, const double* d_src, double* d_dst, int64_t n);
int launch_gather_probe(mf_ctx* c, const double* d_src, double* d_dst, int64_t n);
int launch_flush(mf_ctx* c, double* d_buf, int64_t n);
// cache-flush buffer of mf_time_rhs_cold: 512 MiB, twice the 256 MB last-level cache
constexpr size_t kFlushDoubles = (size_t)1 << 26;
int launch_gather_nodes_i8(mf_ctx* c, const int8_t* d_src, int8_t* d_dst);
int launch_sort_stencil(mf_ctx* c);
int launch_pack_sorted(mf_ctx* c, int target, const double* d_wx, const double* d_wy);
int launch_fill_lap_t(mf_ctx* c, int64_t n, const int64_t* d_src, double* d_w, float* d_w32);
int build_tiles(mf_ctx* c, bool want_b, int rec_b);
// tile-major copies of the operators for the tile kernels (after
// build_tiles and after every operator change)
int ensure_tile_ops(mf_ctx* c);

// rhs
enum StageKind { ST_NONE = 0, ST_FSAL = 1, ST_TWO = 2, ST_THREE = 3 };
struct StageOut {
    int kind = ST_NONE;
    double dt = 0.0;
    const double* uprev = nullptr;
    double* out = nullptr;
    // stepping with the residual viscosity: sums_out applies the next
    // rhs!'s first boundary pass (at (t + out_t_adv) + out_dt_off) to the
    // output and leaves the output's per-block sums in d_stage_psum;
    // sums_in: this rhs!'s input came so (its first pass and sum pass are
    // skipped)
    bool sums_out = false, sums_in = false;
    double out_t_adv = 0.0, out_dt_off = 0.0;
};
int run_rhs(mf_ctx* c, int prec, double* d_u, double* d_du, double dt_off, const StageOut& st);
// parts of one rhs! (for timing single kernels)
enum RhsPart { RHS_BC = 1, RHS_REDUCE = 2, RHS_FLUX = 4, RHS_TRANSPOSE = 8, RHS_IGR = 16, RHS_ALL = 31 };
int run_rhs_parts(mf_ctx* c, int prec, double* d_u, double* d_du, double dt_off, const StageOut& st,
                  int parts);
int run_set_time(mf_ctx* c, double t, int success_iter);
// History callback; with advance_dt != 0 it first ends the step (t += dt,
// success_iter += 1) in the same launch.
int run_history(mf_ctx* c, const double* d_u, int init, double advance_dt = 0.0);
int run_advance_time(mf_ctx* c, double dt);
int run_limiter(mf_ctx* c, double* d_u, double thr, int var);
int run_axpby(mf_ctx* c, int64_t n, double a, const double* x, double b, double* y);
int run_fill(mf_ctx* c, double* p, int64_t n, double v);

}  // namespace mf
