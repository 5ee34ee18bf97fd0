/*
 * mf_oracle.h - CPU restatement (FP64, plain C) of MeshfreeTrixi.jl's RBF-FD
 * semidiscretization path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (libmfrbf.so) never links,
 * loads or calls it.
 *
 * Every entry point cites the reference file:line whose arithmetic it
 * restates.  Compiled with -ffp-contract=off: multiply-adds are fused only
 * where the reference sits inside a `@muladd` block, and there explicitly
 * through fma().
 */
#ifndef MF_ORACLE_H
#define MF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- clouds */
uint64_t or_splitmix64(uint64_t* state);
double or_uniform01(uint64_t* state);
/* s x s jittered unit-square cloud; see the generator comment in the .c */
void or_cloud_square(int s, uint64_t seed, double jitter_frac, double* x, double* y,
                     int8_t* tag, double* nx, double* ny);

/* ------------------------------------------------------------------- kNN */
void or_knn_brute(const double* x, const double* y, int64_t N, int k, int32_t* idx,
                  double* d2);
int or_knn_grid(const double* x, const double* y, int64_t N, int k, int32_t* idx,
                double* d2);
void or_dx_stats(const double* d2_nn, int64_t N, double* dx_min, double* dx_avg);

/* --------------------------------------------------------------- weights */
int or_bk_factor(int m, double* A, int lda, int* ipiv);
void or_bk_solve(int m, const double* A, int lda, const int* ipiv, int nrhs, double* B,
                 int ldb);
int or_npoly(int deg);
int or_weights(const double* x, const double* y, int64_t N, const int64_t* ptr,
               const int32_t* idx, int phs_p, int poly_deg, int deriv_k, double* wx,
               double* wy);
/* or_weights on OpenMP threads (chunks of 64 stencils, the same bits);
 * nthreads <= 0 keeps OpenMP's default */
int or_weights_par(const double* x, const double* y, int64_t N, const int64_t* ptr,
                   const int32_t* idx, int phs_p, int poly_deg, int deriv_k, double* wx,
                   double* wy, int nthreads);
double or_phs_deriv(int p, int k, int dir, double x, double y);

/* ------------------------------------------------------------ sparse CSC */
void or_csc_build(int64_t N, const int64_t* ptr, const int32_t* idx, int64_t* colptr,
                  int32_t* rowval, int64_t* perm);
void or_csc_mul(int64_t N, const int64_t* colptr, const int32_t* rowval, const double* nzval,
                const double* x, double* out, double alpha, int overwrite);
void or_csc_tmul(int64_t N, const int64_t* colptr, const int32_t* rowval,
                 const double* nzval, const double* x, double* out, double alpha);

/* ------------------------------------------------------------------ rhs! */
enum { OR_EQ_ADVECTION = 0, OR_EQ_EULER = 1 };
enum { OR_BC_DIRICHLET = 0, OR_BC_SLIP = 1, OR_BC_NOTHING = 2 };
enum { OR_FN_CONST = 0, OR_FN_PULSE = 1, OR_FN_WAVE = 2, OR_FN_GRADIENT = 3, OR_FN_TABLE = 4 };
enum { OR_RV_NONE = 0, OR_RV_RESIDUAL = 1, OR_RV_UPWIND = 2 };

#define OR_SRC_FLYER 1
#define OR_SRC_TOMINEC 2
#define OR_SRC_VISCOSITY 3
#define OR_SRC_IGR 4

typedef struct {
    int64_t N;
    int nvar;
    int eq;
    double adv_a[2];
    double gamma;
    const double* px;
    const double* py;
    /* CSC operators sharing one pattern */
    const int64_t* colptr;
    const int32_t* rowval;
    const double* dx;
    const double* dy;
    const double* hv; /* NULL = no Flyer hyperviscosity */
    double hv_gamma;
    int rv_mode;
    double c_rv, c_uw, dx_avg;
    int hv_first; /* 1: HV source before RV source (keyword order) */
    /* boundary tags, in NamedTuple order */
    int ntags;
    const int32_t* tag_ptr; /* ntags+1 */
    const int32_t* tag_kind;
    const int32_t* tag_fn;
    const double* tag_fn_params; /* 8 per tag */
    const int32_t* bidx;
    const double* bnormal; /* 2 per boundary entry */
    const double* btable;  /* nvar per boundary entry, OR_FN_TABLE */
    /* residual-viscosity state */
    const double* approx_du;
    int success_iter;
    /* Tominec hyperviscosity: du -= tom_gamma lap' (lap u); NULL = none */
    const double* lap;
    double tom_gamma;
    /* sources in keyword order (OR_SRC_*), 0-terminated; src_order[0] = 0:
     * Flyer and the viscosity source ordered by hv_first, then Tominec,
     * then IGR */
    int src_order[4];
    /* IGR source (Euler): alpha, CG iterations; igr_alpha = 0: none.
     * igr_sigma_out (N doubles, may be NULL) receives sigma. */
    double igr_alpha;
    int igr_maxiter;
    double* igr_sigma_out;
} or_rhs_cfg;

void or_flux(int eq, const double* adv_a, double gamma, int orientation, const double* u,
             double* f);
void or_bc_value(int fn, const double* params, int nvar, double gamma, double x, double y,
                 double t, double* out);
void or_bc_pass(const or_rhs_cfg* c, double t, double* u, double* du);
int or_rhs(const or_rhs_cfg* c, double t, double* u, double* du, double* eps_out,
           int8_t* eps_c_out, double* scratch);
int64_t or_rhs_scratch_len(const or_rhs_cfg* c);

/* IGR pieces (sources/IGR.jl).  rinv = 1 / rho.  or_igr_rhs: b and rinv
 * from the state (update_igr_rhs!, :117-157, and :176-178);
 * or_igr_apply: y = (rinv - alpha (Dx rinv Dx + Dy rinv Dy)) v (the
 * IGRCompositeMatrix mul!, :50-66); or_igr_cg: IterativeSolvers.jl's cg!
 * from x (in/out), returns the iterations done. */
void or_igr_rhs(const or_rhs_cfg* c, const double* u, double* b, double* rinv);
void or_igr_apply(const or_rhs_cfg* c, const double* rinv, const double* v, double* y);
int or_igr_cg(const or_rhs_cfg* c, const double* rinv, const double* b, double* x, int maxiter);

/* residual viscosity pieces (hyperviscosity.jl) */
void or_upwind_visc(const or_rhs_cfg* c, const double* u, double* eps_uw);
void or_rv_norms(int64_t N, int nvar, const double* u, double* mean, double* ninf);
void or_residual_visc(const or_rhs_cfg* c, const double* du, const double* u, double* eps_rv);
void or_update_visc(int64_t N, const double* eps_uw, const double* eps_rv, int success_iter,
                    double* eps, int8_t* eps_c);
double or_pairwise_sum(const double* a, int64_t lo, int64_t hi);

/* history callback (history.jl) */
void or_time_deriv_weights(int q, const double* t, double* w);
void or_history_update(int64_t nvals, int approx_order, int success_iter, double t,
                       const double* u, double* time_hist, double* sol_hist,
                       double* approx_du);

/* Zhang-Shu stage limiter (positivity_zhang_shu_point2d.jl) */
enum { OR_VAR_DENSITY = 0, OR_VAR_PRESSURE = 1 };
void or_limiter_zhang_shu(int64_t N, int nvar, double gamma, const int64_t* ptr,
                          const int32_t* idx, double threshold, int variable, double* u,
                          double* scratch);

/* Multi-threaded (OpenMP) rhs! for the CPU baseline timing: Euler with
 * the residual or upwind viscosity and strong BCs only (-1 otherwise);
 * row-form stencil rptr/ridx/rwx/rwy; agrees with or_rhs to rounding.
 * scratch: 17 N doubles; nthreads <= 0 keeps OpenMP's default. */
int or_rhs_par(const or_rhs_cfg* c, const int64_t* rptr, const int32_t* ridx, const double* rwx,
               const double* rwy, double t, double* u, double* du, double* scratch, int nthreads);

/* SSPRK33 with FSAL, history callback and optional stage limiter */
typedef struct {
    int approx_order; /* 0 = no history callback */
    int limiter_on;
    double lim_threshold[2];
    int lim_variable[2];
    int lim_count;
    const int64_t* ptr; /* stencil for the limiter */
    const int32_t* idx;
} or_step_cfg;
int or_ssprk33(or_rhs_cfg* c, const or_step_cfg* s, double t0, double dt, int nsteps,
               double* u, double* fsal, double* time_hist, double* sol_hist,
               double* approx_du, int* success_iter_io);

/* adjoint: vector-Jacobian product of rhs! (Euler/advection, BCs, no sources) */
int or_rhs_vjp(const or_rhs_cfg* c, double t, const double* u_in, const double* lam,
               double* ubar, double* scratch);

#ifdef __cplusplus
}
#endif
#endif
