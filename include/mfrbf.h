/*
 * mfrbf.h - C ABI of libmfrbf.so, the MI355X (gfx950) implementation of the
 * RBF-FD semidiscretization path of MeshfreeTrixi.jl.
 *
 * Conventions
 *  - Every function returns an int status: MF_OK (0) or one of MF_ERR_*;
 *    mf_last_error() then describes the failure.  Nothing throws.
 *  - Host pointers are read or written only during the call.
 *  - State vectors cross the ABI as nvar separate host arrays of n_points
 *    doubles ("fields", the reference's StructArray components,
 *    pointcloudsolver/rbfsolver.jl:111-116).  Inside the context they live
 *    in device memory, node-major (all variables of a node contiguous).
 *  - Node and stencil indices are 0-based (the reference's are 1-based).
 *  - One context = one device + one HIP stream.  A context is not
 *    re-entrant; distinct contexts may be used from different threads.
 *
 * Each declaration names the reference interface (file:line, relative to
 * the MeshfreeTrixi.jl tree) that it replaces.
 */
#ifndef MFRBF_H
#define MFRBF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mf_ctx mf_ctx;

/* status codes */
#define MF_OK 0
#define MF_ERR_ARGUMENT 1    /* invalid argument (null pointer, size, enum) */
#define MF_ERR_HIP 2         /* HIP runtime error */
#define MF_ERR_STATE 3       /* call out of order (e.g. no stencil yet) */
#define MF_ERR_SINGULAR 4    /* a stencil's saddle-point system is singular */
#define MF_ERR_UNSUPPORTED 5 /* combination not provided by this build */

/* equations (Trixi's CompressibleEulerEquations2D / LinearScalarAdvectionEquation2D) */
#define MF_EQ_ADVECTION_2D 0
#define MF_EQ_EULER_2D 1

/* boundary condition kinds, PointCloudBCs.jl */
#define MF_BC_DIRICHLET 0 /* BoundaryConditionDirichlet, :49-63 */
#define MF_BC_SLIP_WALL 1 /* boundary_condition_slip_wall, :87-106 */
#define MF_BC_DO_NOTHING 2 /* BoundaryConditionDoNothing, :108-115 */

/* built-in Dirichlet value functions g(x, y, t); parameters, 8 per tag:
 *  CONST     params[0..nvar)
 *  PULSE     scalar exp(-((x - a1 t - x0)^2 + (y - a2 t - y0)^2) / w^2),
 *            params (x0, y0, w, a1, a2)
 *  WAVE      Euler density wave rho = r0 + A sin(2 pi (x + y - t (v1 + v2)) + phi),
 *            velocity (v1, v2), pressure p; params (r0, A, v1, v2, p, phi)
 *  GRADIENT  Euler linear state of the reference's operator tests
 *  TABLE     values uploaded with mf_set_boundary_values */
#define MF_FN_CONST 0
#define MF_FN_PULSE 1
#define MF_FN_WAVE 2
#define MF_FN_GRADIENT 3
#define MF_FN_TABLE 4

/* operators */
#define MF_OP_FLUX 0            /* Dx, Dy (compute_operators.jl:409-453) */
#define MF_OP_HYPERVISCOSITY 1  /* Dx^2k + Dy^2k (hyperviscosity.jl:43-44) */
#define MF_OP_LAPLACIAN 2       /* lap = Dxx + Dyy of the Tominec source (hyperviscosity.jl:111-113) */

/* sources, for mf_set_source_order (SourceTerms keyword order) */
#define MF_SRC_FLYER 1     /* SourceHyperviscosityFlyer */
#define MF_SRC_TOMINEC 2   /* SourceHyperviscosityTominec */
#define MF_SRC_VISCOSITY 3 /* SourceResidual/UpwindViscosityTominec */
#define MF_SRC_IGR 4       /* SourceIGR */

/* viscosity sources (hyperviscosity.jl) */
#define MF_VISC_NONE 0
#define MF_VISC_RESIDUAL 1 /* SourceResidualViscosityTominec, :181-244, :382-409 */
#define MF_VISC_UPWIND 2   /* SourceUpwindViscosityTominec, :148-167, :351-380 */

/* RHS precision */
#define MF_PREC_F64 0     /* FP64 weights, state and accumulation */
#define MF_PREC_SPLIT32 1 /* FP32 weights, centred FP64->FP32 differences,
                             FP32 accumulation; FP64 state */
#define MF_PREC_F64_REF 2 /* FP64 in the reference's operation order: per
                             node, terms in ascending column order, products
                             rounded before each add, x terms then y terms
                             then sources (the order of SparseArrays' CSC
                             mul!), reproducible bit for bit */

/* state buffers owned by a context */
#define MF_BUF_U 0      /* integrator state u */
#define MF_BUF_DU 1     /* rhs output du */
#define MF_BUF_AUX0 2   /* caller scratch (e.g. adjoint seed) */
#define MF_BUF_AUX1 3   /* caller scratch (e.g. adjoint result) */

/* Zhang-Shu limiter variables (Trixi.density, Trixi.pressure) */
#define MF_VAR_DENSITY 0
#define MF_VAR_PRESSURE 1

/* mf_ssprk33 flags */
#define MF_STEP_GRAPH 1   /* capture one step as a hipGraph and replay it */
#define MF_STEP_LIMITER 2 /* apply the configured stage limiter */
#define MF_STEP_HISTORY 4 /* HistoryCallback at init and after every step */
#define MF_STEP_NOSYNC 8  /* return once the steps are enqueued on the context's stream
                             (mf_synchronize before touching its buffers); lets one host
                             thread keep several contexts (an ensemble) busy at once */

/* ------------------------------------------------------------- context */
/* Creates a context on HIP device `device` with its own stream. */
int mf_ctx_create(int device, mf_ctx** out);
int mf_ctx_destroy(mf_ctx* ctx);
/* Message of the last failed call on ctx ("" if none); valid until the next call. */
const char* mf_last_error(const mf_ctx* ctx);
/* Blocks until all work queued on ctx's stream has finished. */
int mf_synchronize(mf_ctx* ctx);

/* --------------------------------------------------------------- cloud */
/* Node coordinates.  Replaces the `points` of PointData
 * (domains/PointCloudDomain/geometry_primatives.jl:297-304, :322-339).
 * MF_ERR_ARGUMENT for fewer than 2 points or a non-finite coordinate. */
int mf_set_points(mf_ctx* ctx, const double* x, const double* y, int64_t n_points);

/* The same for a point set made of n_groups independent clouds (an
 * ensemble of runs in one context): group g is points group_ptr[g] ..
 * group_ptr[g+1] (group_ptr[0] = 0, group_ptr[n_groups] = n_points).  The
 * internal node order keeps each group contiguous (Morton order inside the
 * group), so every kernel launch of rhs! and of the SSPRK33 step covers all
 * clouds while no LDS tile mixes more than two of them.  Stencils must stay
 * inside their group: mf_set_stencil returns MF_ERR_ARGUMENT for one that
 * crosses, and mf_knn / mf_knn_noalloc return MF_ERR_UNSUPPORTED on a
 * grouped set (search each cloud on its own, as a set of its own).  The
 * residual viscosity and IGR sources, which reduce over all points, return
 * MF_ERR_UNSUPPORTED at rhs! on a grouped set.  No counterpart in the
 * reference (which runs one cloud per semidiscretization). */
int mf_set_points_grouped(mf_ctx* ctx, const double* x, const double* y, int64_t n_points, int n_groups,
                          const int64_t* group_ptr);

/* k nearest neighbours of every node (self first), ordered by squared
 * distance, ties by index.  Replaces knn(KDTree(points), points, nv, true)
 * and the dx_min/dx_avg statistics of PointData(points, basis)
 * (geometry_primatives.jl:327-334).  The stencil stays on the device for
 * mf_compute_operator.  idx_out (n_points*k, row per node) and dist_out
 * (Euclidean distances, same shape) may be NULL. */
int mf_knn(mf_ctx* ctx, int k, int32_t* idx_out, double* dist_out, double* dx_min,
           double* dx_avg);

/* mf_knn without allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  No counterpart in
 * the reference.  The same stencils as mf_knn (self first, ordered by
 * squared distance, ties by index) for the points of mf_set_points, written
 * to caller-owned DEVICE memory: d_idx_out (n_points*k int32, row per node,
 * the caller's indices) and d_dist_out (Euclidean distances, same layout;
 * may be NULL).  Unlike mf_knn it leaves the context's stencil unchanged.
 * The search's intermediates live in a 9 MiB static device array that each
 * GPU holds until the process ends: at most 262,144 points (and as many
 * search-grid cells); larger clouds return MF_ERR_UNSUPPORTED.  Every call
 * on one GPU shares that array, so calls that overlap on one GPU (from
 * other threads or contexts) race. */
int mf_knn_noalloc(mf_ctx* ctx, int k, int32_t* d_idx_out, double* d_dist_out);

/* Uploads a stencil computed elsewhere (n_points x k, row per node, self
 * first); replaces PointData.neighbors (geometry_primatives.jl:299). */
int mf_set_stencil(mf_ctx* ctx, const int32_t* idx, int64_t n_points, int k);

/* ------------------------------------------------------------- weights */
/* RBF-FD weights on the current stencil: PHS r^phs_power plus monomials of
 * total degree <= poly_degree, derivative order deriv_order along x and y.
 * target MF_OP_FLUX (deriv_order 1) replaces compute_flux_operator(solver,
 * domain) (pointcloudsolver/compute_operators.jl:409-453); target
 * MF_OP_HYPERVISCOSITY (deriv_order 2k) replaces compute_flux_operator(
 * solver, domain, 2k) summed into one operator (compute_operators.jl:549-594,
 * sources/hyperviscosity.jl:43-44); target MF_OP_LAPLACIAN (deriv_order 2)
 * replaces the lap = sum(compute_flux_operator(solver, domain, 2)) of the
 * Tominec source (hyperviscosity.jl:111-113).  Weights stay on the device; wx_out /
 * wy_out (n_points*k, row per node, may be NULL) receive the x / y operator.
 * Returns MF_ERR_SINGULAR if some stencil's system is singular. */
int mf_compute_operator(mf_ctx* ctx, int phs_power, int poly_degree, int deriv_order,
                        int target, double* wx_out, double* wy_out);

/* Installs host weights for the current stencil (n_points*k each, row per
 * node).  MF_OP_FLUX: wx = Dx, wy = Dy.  MF_OP_HYPERVISCOSITY: wx = H,
 * MF_OP_LAPLACIAN: wx = lap; wy must be NULL for both.  Replaces the cached rbf_differentiation_matrices /
 * hv_differentiation_matrix (pointcloudsolver/rbfsolver.jl:139,162). */
int mf_set_operator(mf_ctx* ctx, int target, const double* wx, const double* wy);

/* Batched weights for ragged stencils over the cloud of mf_set_points:
 * stencil e = idx[ptr[e] .. ptr[e+1]), centre first, 1 to 64 nodes each.
 * The same per-stencil system as compute_operators.jl:422-441 (the
 * reference's per-node loop, M \ rhs at :435), for many stencils of uneven
 * size.  By default (MF_TUNE_WEIGHTS_SOLVER 0 or 2) the saddle systems are
 * solved in their null-space form on the FP64 matrix cores, one workgroup
 * per stencil in one launch per size variant (stencils of up to 32, of 33
 * to 48 and of 49 to 64 nodes; odd PHS power p, polynomial degree from
 * ceil(p/2) - 1 (where r^p is conditionally positive definite) to 4);
 * stencils of fewer than twice the monomial count, every stencil of another
 * basis, and all of them with MF_TUNE_WEIGHTS_SOLVER 1, by the
 * Bunch-Kaufman restatement, one launch per stencil size.  ptr/idx are
 * uploaded once; mf_weights_ragged_run recomputes on the device (timing
 * loop), mf_weights_ragged_get copies the result (layout like idx). */
int mf_weights_ragged_set(mf_ctx* ctx, const int64_t* ptr, const int32_t* idx,
                          int64_t n_stencils);
int mf_weights_ragged_run(mf_ctx* ctx, int phs_power, int poly_degree, int deriv_order);
int mf_weights_ragged_get(mf_ctx* ctx, double* wx_out, double* wy_out);

/* ------------------------------------------------------------- physics */
/* MF_EQ_ADVECTION_2D: p0, p1 = advection velocity; MF_EQ_EULER_2D: p0 = gamma.
 * Replaces the `equations` argument of Trixi.rhs! (rbfsolver.jl:397). */
int mf_set_equations(mf_ctx* ctx, int equations, double p0, double p1);

/* Boundary conditions, one tag after another in the order of the
 * reference's boundary_conditions NamedTuple (rbfsolver.jl:277-286).
 * Tag g covers entries tag_ptr[g] .. tag_ptr[g+1] of node_idx / normals
 * (2 per entry); tag_kind[g] is MF_BC_*, tag_fn[g] the MF_FN_* value
 * function of a Dirichlet tag, tag_params 8 doubles per tag.  Replaces
 * domain.boundary_tags + boundary_conditions (rbfsolver.jl:288-318). */
int mf_set_boundary(mf_ctx* ctx, int n_tags, const int32_t* tag_ptr, const int32_t* tag_kind,
                    const int32_t* tag_fn, const double* tag_params, const int32_t* node_idx,
                    const double* normals);
/* Values for MF_FN_TABLE tags: nvar doubles per boundary entry. */
int mf_set_boundary_values(mf_ctx* ctx, const double* values);

/* Flyer hyperviscosity: du -= gamma * H u with the MF_OP_HYPERVISCOSITY
 * operator; gamma = c * dx_min^(2k).  Replaces SourceHyperviscosityFlyer
 * (hyperviscosity.jl:14-64).  gamma = 0 removes the source. */
int mf_set_source_hyperviscosity(mf_ctx* ctx, double gamma);

/* Residual / upwind viscosity: du -= sum_d D_d^T (eps .* (D_d u)).
 * Replaces SourceResidualViscosityTominec / SourceUpwindViscosityTominec
 * (hyperviscosity.jl:181-409).  h = dx_avg.  hv_first = 1 if the
 * hyperviscosity source comes before this one in the SourceTerms keyword
 * order (generic_sources.jl:7-18, rbfsolver.jl:388-395). */
int mf_set_source_viscosity(mf_ctx* ctx, int mode, double c_rv, double c_uw, double h,
                            int approx_order, int hv_first);

/* Tominec hyperviscosity: du -= gamma * lap' (lap u) with the
 * MF_OP_LAPLACIAN operator; gamma = c * dx_min^4.5.  Replaces
 * SourceHyperviscosityTominec (hyperviscosity.jl:66-134), which forms
 * H = lap' * lap explicitly and applies du += H (u * -gamma); here H is
 * applied as its two factors (the same operator; rounding differs).
 * gamma = 0 removes the source.  Euler equations only. */
int mf_set_source_hyperviscosity_tominec(mf_ctx* ctx, double gamma);

/* The order in which the sources add to du, as the keys of the reference's
 * SourceTerms NamedTuple (generic_sources.jl:7-18, rbfsolver.jl:388-395):
 * kinds[0..n) is a list of distinct MF_SRC_* values; sources not listed
 * follow in the order FLYER, TOMINEC, VISCOSITY.  It decides what the
 * residual viscosity sees in du and, for MF_PREC_F64_REF, the rounding
 * order.  A Tominec source before a residual viscosity source is
 * MF_ERR_UNSUPPORTED at the next rhs! (its du would need lap'(lap u)
 * before the viscosity's gathers).  mf_set_source_viscosity's hv_first
 * sets the order of FLYER and VISCOSITY only. */
int mf_set_source_order(mf_ctx* ctx, const int32_t* kinds, int n);

/* IGR source (Euler equations): sigma from maxiter conjugate-gradient
 * iterations, started at 0, on (1/rho - alpha (Dx diag(1/rho) Dx + Dy
 * diag(1/rho) Dy)) sigma = alpha (trace^2 + trace_sq), then du_m1 -= Dx
 * sigma, du_m2 -= Dy sigma.  Replaces SourceIGR (sources/IGR.jl:17-239)
 * with its CG solve (IterativeSolvers cg! semantics, maxiter = 20 in the
 * reference); trace = D_x rho + D_y v1 as the reference computes it.  It
 * must come after the other sources in the keyword order
 * (MF_ERR_UNSUPPORTED otherwise).  The CG runs in FP64 in every precision.
 * alpha = 0 removes the source. */
int mf_set_source_igr(mf_ctx* ctx, double alpha, int maxiter);
/* sigma of the last rhs! with the IGR source (n_points doubles, the
 * caller's order); replaces SourceIGR's cache.sigma (IGR.jl:113), which
 * update_sigma! fills (IGR.jl:170-190). */
int mf_get_igr_sigma(mf_ctx* ctx, double* sigma_out);

/* --------------------------------------------------------------- state */
/* Copies nvar host fields into / out of a context buffer (MF_BUF_*). */
int mf_state_upload(mf_ctx* ctx, int buf, const double* const* fields);
int mf_state_download(mf_ctx* ctx, int buf, double* const* fields);
/* Device-side copy of one context buffer into another (e.g. keep a state
 * while stepping on); a plain streaming copy of nvar * n_points doubles.
 * No counterpart in the reference (copyto! on host arrays there). */
int mf_state_copy(mf_ctx* ctx, int dst_buf, int src_buf);
/* A known-bytes kernel from buffer src_buf into dst_buf (nvar * n_points
 * doubles, each read once and written once: 16 bytes per double): kind 0
 * the streaming copy of mf_state_copy, kind 1 a gather whose wavefronts
 * read scattered doubles within 2 KiB windows (the access shape of a
 * stencil gather); kind 2 ignores the buffers and streams a read-modify-
 * write over a 512 MiB scratch array (1 GiB of HBM traffic, twice the
 * last-level cache: the device's streaming bandwidth).  ms (may be NULL)
 * receives the kernel's time from HIP events.  No counterpart in the
 * reference; the calibration kernels of the HBM traffic counters and a
 * practical bandwidth ceiling. */
int mf_bandwidth_probe(mf_ctx* ctx, int kind, int dst_buf, int src_buf, double* ms);
/* Device address of a context buffer: node-major, nvar doubles per node,
 * nodes in the context's internal order (see mf_node_order). */
int mf_state_device_ptr(mf_ctx* ctx, int buf, double** dptr);
/* The internal node order: order_out[p] = the caller's index of internal
 * node p (n_points int32).  The context stores nodes along a Morton curve
 * so that stencil neighbours are close in memory; every other entry point
 * takes and returns data in the caller's order.  No counterpart in the
 * reference (its arrays are in the caller's order); needed only with
 * mf_state_device_ptr. */
int mf_node_order(mf_ctx* ctx, int32_t* order_out);

/* ----------------------------------------------------------------- rhs! */
/* du = rhs(u, t) on device buffers; u (MF_BUF_U) is mutated at boundary
 * nodes exactly like the reference.  Replaces Trixi.rhs!(du, u, t, domain,
 * equations, ic, bcs, sources, solver, cache), rbfsolver.jl:397-428. */
int mf_rhs(mf_ctx* ctx, double t, int precision);

/* The same on host fields (upload, rhs!, download of du and u). */
int mf_rhs_host(mf_ctx* ctx, double t, int precision, double* const* u, double* const* du);

/* Viscosity of the last rhs! with a viscosity source: eps and the choice
 * code eps_c (0 residual, 1 upwind, 2 eps()), hyperviscosity.jl:331-349. */
int mf_get_viscosity(mf_ctx* ctx, double* eps, int8_t* eps_c);

/* Inspection (no counterpart in the reference, where these are locals of
 * update_residual_visc!, hyperviscosity.jl:289-311): mean(u) and ninf, the
 * lexicographic maximum of |u - mean| with zeros replaced by eps(), of the
 * last rhs! with the residual viscosity.  Either pointer may be NULL. */
int mf_get_viscosity_norms(mf_ctx* ctx, double* mean, double* ninf);

/* --------------------------------------------------- step/stage hooks */
/* HistoryCallback affect! for the viscosity source: shift the solution
 * history with u = MF_BUF_U at time t and rebuild approx_du
 * (modify_cache!, callbacks_step/history.jl:91-103; shift_soln_history!,
 * update_approx_du!, time_deriv_weights!, :105-152).  success_iter =
 * integrator.success_iter. */
int mf_history_update(mf_ctx* ctx, double t, int success_iter);
int mf_get_approx_du(mf_ctx* ctx, double* const* fields);

/* Stage limiter: Trixi.limiter_zhang_shu!(u, threshold, variable, domain,
 * ...) on MF_BUF_U (callbacks_stage/positivity_zhang_shu_point2d.jl:22-83). */
int mf_limiter_zhang_shu(mf_ctx* ctx, double threshold, int variable);
/* Limiter used by mf_ssprk33 with MF_STEP_LIMITER (up to 2 variables). */
int mf_set_stage_limiter(mf_ctx* ctx, int n, const double* thresholds, const int32_t* variables);

/* ------------------------------------------------------ time stepping */
/* n_steps of SSPRK33 (Shu-Osher form, FSAL) from MF_BUF_U at t0 with step
 * dt; with MF_STEP_HISTORY and a residual viscosity source the
 * HistoryCallback runs at init (when resume == 0) and after every step.
 * Replaces solve(ode, SSPRK33(...), dt, callbacks) on the reference's
 * semidiscretization.  resume = 1 continues from the previous call's FSAL
 * and history state (same dt; MF_BUF_U as that call left it).  With the
 * residual viscosity and no stage limiter each stage is written with the
 * next rhs!'s first boundary pass and sums already applied, so a resume
 * that switches MF_STEP_LIMITER on or off there (or the residual viscosity
 * on or off) returns MF_ERR_STATE. */
int mf_ssprk33(mf_ctx* ctx, double t0, double dt, int n_steps, int precision, int flags,
               int resume);

/* ------------------------------------------------------------- adjoint */
/* ubar = (d rhs / d u)^T lambda at u = MF_BUF_U (flux divergence, boundary
 * conditions; no sources) with lambda = MF_BUF_AUX0, result in MF_BUF_AUX1.
 * No counterpart in the reference (which has no adjoint). */
int mf_rhs_vjp(mf_ctx* ctx, double t);

/* Gradient of L = sum_i rho_i(T)^2 after n_steps SSPRK33 steps from
 * u0 = MF_BUF_U with respect to u0 (result in MF_BUF_AUX1, all
 * variables); checkpoints every `checkpoint_every` steps.  No counterpart
 * in the reference. */
int mf_ssprk33_vjp(mf_ctx* ctx, double t0, double dt, int n_steps, int checkpoint_every,
                   double* loss);

/* ------------------------------------------------------------- tuning */
#define MF_TUNE_FLUX_BATCH 0      /* stencil entries loaded per batch: 3, 5, 6, 10 (0 = default) */
#define MF_TUNE_TRANSPOSE_BATCH 1 /* transposed entries loaded per batch: 2, 4, 8 (0 = default) */
#define MF_TUNE_NODE_ORDER 2      /* internal node order for the next mf_set_points: 0 Morton
                                     (default), 1 the caller's order; precision is ignored */
#define MF_TUNE_TILES 3           /* 1 (default): stage stencil halos in LDS where they fit;
                                     0: gather from global memory; precision is ignored */
#define MF_TUNE_WEIGHTS_SOLVER 4  /* saddle-system solver of the weights: 0 (default) the
                                     null-space form on the matrix cores for ragged batches
                                     and Bunch-Kaufman (the reference's factorization) for
                                     mf_compute_operator; 1 Bunch-Kaufman for both; 2 the
                                     null-space form for both (where the basis allows it
                                     and stencils have at least twice as many nodes as
                                     monomials); precision is ignored */
/* Selects a compiled kernel variant for one precision (MF_PREC_*, or -1 for
 * all).  A batch size that does not divide the stencil size falls back to
 * the default.  Results do not depend on it (the per-node operation order is
 * the same for every batch size).  No counterpart in the reference (the CPU
 * path has no launch shapes); set by the autotuner of the Python package. */
int mf_set_tuning(mf_ctx* ctx, int key, int precision, int value);

/* State of the LDS tiles after the last rhs! preparation: info[0] phase A
 * tiles in use, [1] phase B tiles in use, [2] tiles, [3] total phase-A halo
 * entries, [4] total phase-B halo entries, [5] largest phase-A halo, [6]
 * largest phase-B halo, [7] phase-B entries including padding (8 int64).
 * No counterpart in the reference; for performance accounting. */
int mf_tile_info(mf_ctx* ctx, int64_t* info);

/* ------------------------------------------------------------- timing */
/* Average device time of one rhs! (ms) over n_calls back-to-back calls on
 * MF_BUF_U, measured with HIP events on the context stream (chains of the
 * captured call).  kernel_ms (4 doubles, may be NULL) receives each part's
 * average kernel time inside n_calls further whole calls (the dispatches
 * carry start/stop events, so each kernel is timed alone while it sees the
 * caches as a full call leaves them):
 * [0] the stencil kernel (flux divergence, sources, phase A of the
 * viscosity), [1] the transposed gather of the viscosity, [2] the two
 * global reductions of the residual viscosity (with the first boundary pass
 * when it rides on them), [3] the boundary passes. */
int mf_time_rhs(mf_ctx* ctx, double t, int precision, int n_calls, double* rhs_ms,
                double* kernel_ms);

/* Median device time (ms) of one eager rhs! on MF_BUF_U over n_calls calls,
 * each after a read-modify-write of a 512 MiB buffer (allocated by the first
 * call, kept until mf_ctx_destroy) that leaves none of the call's data in
 * the L2 or last-level caches: the cold-cache counterpart of mf_time_rhs.
 * No counterpart in the reference; for performance accounting. */
int mf_time_rhs_cold(mf_ctx* ctx, double t, int precision, int n_calls, double* rhs_ms);

#ifdef __cplusplus
}
#endif
#endif
