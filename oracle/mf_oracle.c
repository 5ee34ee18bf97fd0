/*
 * mf_oracle.c - CPU restatement (FP64, plain C) of MeshfreeTrixi.jl's RBF-FD
 * semidiscretization path.  TEST INFRASTRUCTURE ONLY (see mf_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).  Reference files are
 * cited relative to the MeshfreeTrixi.jl source tree.
 *
 * Arithmetic that the reference delegates to packages absent from this
 * tree (NearestNeighbors, LAPACK, SparseArrays, Trixi, OrdinaryDiffEq) is
 * restated from their published algorithms; each such place says so.
 */
#include <omp.h>
#include "mf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Clouds.  Workload generator for the benchmark configurations; it has no
 * counterpart in the reference (whose tests read a Medusa file instead).
 * splitmix64 (Steele, Lea, Flood 2014) drives every draw.
 * ==================================================================== */
uint64_t or_splitmix64(uint64_t* state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double or_uniform01(uint64_t* state)
{
    return (double)(or_splitmix64(state) >> 11) * 0x1.0p-53;
}

/* s x s lattice on [0,1]^2, node k = j*s + i at (i/(s-1), j/(s-1)).
 * Interior nodes: both coordinates jittered by U(-a, a), a = jitter_frac*h
 * (x drawn before y).  Edge nodes: jittered along the edge only.  Corners
 * fixed.  tag: 0 interior, 1 inflow (x = 0 or y = 0), 2 outflow (rest of
 * the boundary).  Outward face normals; corners get the diagonal. */
void or_cloud_square(int s, uint64_t seed, double jitter_frac, double* x, double* y,
                     int8_t* tag, double* nx, double* ny)
{
    uint64_t st = seed;
    const double h = 1.0 / (double)(s - 1);
    const double a = jitter_frac * h;
    const double r2 = 1.0 / sqrt(2.0);
    for (int j = 0; j < s; ++j) {
        for (int i = 0; i < s; ++i) {
            const int64_t k = (int64_t)j * s + i;
            const double bx = (double)i / (double)(s - 1);
            const double by = (double)j / (double)(s - 1);
            const int bl = (i == 0), br = (i == s - 1), bb = (j == 0), bt = (j == s - 1);
            const int vert = bl || br, horz = bb || bt;
            double px = bx, py = by;
            if (vert && horz) {
                /* corner: fixed */
            } else if (vert) {
                double u = or_uniform01(&st);
                py = by + (2.0 * u - 1.0) * a;
            } else if (horz) {
                double u = or_uniform01(&st);
                px = bx + (2.0 * u - 1.0) * a;
            } else {
                double ux = or_uniform01(&st);
                double uy = or_uniform01(&st);
                px = bx + (2.0 * ux - 1.0) * a;
                py = by + (2.0 * uy - 1.0) * a;
            }
            x[k] = px;
            y[k] = py;
            int8_t tg = 0;
            double nnx = 0.0, nny = 0.0;
            if (vert || horz) {
                tg = (bl || bb) ? 1 : 2;
                if (vert && horz) {
                    nnx = br ? r2 : -r2;
                    nny = bt ? r2 : -r2;
                } else if (bl) {
                    nnx = -1.0;
                } else if (br) {
                    nnx = 1.0;
                } else if (bb) {
                    nny = -1.0;
                } else {
                    nny = 1.0;
                }
            }
            if (tag) tag[k] = tg;
            if (nx) nx[k] = nnx;
            if (ny) ny[k] = nny;
        }
    }
}

/* ======================================================================
 * kNN.  Reference: PointData(points, basis), geometry_primatives.jl:322-339,
 * which calls NearestNeighbors' knn(KDTree(points), points, nv, true).
 * NearestNeighbors compares reduced (squared Euclidean) distances
 *   d2 = (xi - xj)^2 + (yi - yj)^2   (no fused multiply-add)
 * and sorts the result by distance.  Ties have no defined order there; here
 * they are broken by index, so (d2, idx) is a total order.
 * ==================================================================== */
static inline double or_d2(double xi, double yi, double xj, double yj)
{
    double dx = xi - xj, dy = yi - yj;
    return dx * dx + dy * dy;
}

static inline int or_before(double d2a, int32_t ia, double d2b, int32_t ib)
{
    return (d2a < d2b) || (d2a == d2b && ia < ib);
}

/* insert (d2, j) into the sorted top-k list (length *cnt <= k) */
static inline void or_topk_insert(double* bd, int32_t* bi, int* cnt, int k, double d2, int32_t j)
{
    int c = *cnt;
    if (c == k && !or_before(d2, j, bd[k - 1], bi[k - 1])) return;
    int pos = (c < k) ? c : k - 1;
    while (pos > 0 && or_before(d2, j, bd[pos - 1], bi[pos - 1])) {
        bd[pos] = bd[pos - 1];
        bi[pos] = bi[pos - 1];
        --pos;
    }
    bd[pos] = d2;
    bi[pos] = j;
    if (c < k) *cnt = c + 1;
}

void or_knn_brute(const double* x, const double* y, int64_t N, int k, int32_t* idx, double* d2)
{
    double* bd = (double*)malloc(sizeof(double) * k);
    int32_t* bi = (int32_t*)malloc(sizeof(int32_t) * k);
    for (int64_t i = 0; i < N; ++i) {
        int cnt = 0;
        for (int64_t j = 0; j < N; ++j)
            or_topk_insert(bd, bi, &cnt, k, or_d2(x[i], y[i], x[j], y[j]), (int32_t)j);
        for (int s = 0; s < k; ++s) {
            idx[i * k + s] = bi[s];
            if (d2) d2[i * k + s] = bd[s];
        }
    }
    free(bd);
    free(bi);
}

/* Exact kNN over a uniform grid (same (d2, idx) order as the brute force),
 * for clouds too large for O(N^2).  Rings of cells are searched until the
 * k-th candidate is closer than every cell not yet visited. */
int or_knn_grid(const double* x, const double* y, int64_t N, int k, int32_t* idx, double* d2)
{
    if (N < k || k < 1) return 1;
    double x0 = x[0], x1 = x[0], y0 = y[0], y1 = y[0];
    for (int64_t i = 1; i < N; ++i) {
        if (x[i] < x0) x0 = x[i];
        if (x[i] > x1) x1 = x[i];
        if (y[i] < y0) y0 = y[i];
        if (y[i] > y1) y1 = y[i];
    }
    /* cells of about two points; their side at least the longer extent / N,
     * so that a flat or elongated cloud gets at most N + 1 cells along its
     * length (the search is exact on any grid) */
    const double W = fmax(x1 - x0, 0.0), H = fmax(y1 - y0, 0.0);
    double cs = fmax(sqrt(2.0 * W * H / (double)N), fmax(W, H) / (double)N);
    if (!(cs > 0)) cs = 1.0;
    int64_t ncx = (int64_t)fmin(floor(W / cs), (double)N) + 1, ncy = (int64_t)fmin(floor(H / cs), (double)N) + 1;
    int64_t ncell = ncx * ncy;
    int64_t* start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * N);
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * N);
    for (int64_t i = 0; i < N; ++i) {
        int64_t cx = (int64_t)floor((x[i] - x0) / cs), cy = (int64_t)floor((y[i] - y0) / cs);
        if (cx < 0) cx = 0;
        if (cx >= ncx) cx = ncx - 1;
        if (cy < 0) cy = 0;
        if (cy >= ncy) cy = ncy - 1;
        cell[i] = cy * ncx + cx;
        start[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ncell);
    memcpy(fill, start, sizeof(int64_t) * ncell);
    for (int64_t i = 0; i < N; ++i) order[fill[cell[i]]++] = (int32_t)i;
    double* bd = (double*)malloc(sizeof(double) * k);
    int32_t* bi = (int32_t*)malloc(sizeof(int32_t) * k);
    for (int64_t i = 0; i < N; ++i) {
        const double qx = x[i], qy = y[i];
        int64_t cx = cell[i] % ncx, cy = cell[i] / ncx;
        int cnt = 0;
        for (int64_t r = 0;; ++r) {
            int64_t ax = cx - r, bx = cx + r, ay = cy - r, by = cy + r;
            for (int64_t jy = ay; jy <= by; ++jy) {
                if (jy < 0 || jy >= ncy) continue;
                for (int64_t jx = ax; jx <= bx; ++jx) {
                    if (jx < 0 || jx >= ncx) continue;
                    if (r > 0 && jy != ay && jy != by && jx != ax && jx != bx) continue;
                    int64_t c = jy * ncx + jx;
                    for (int64_t p = start[c]; p < start[c + 1]; ++p) {
                        int32_t j = order[p];
                        or_topk_insert(bd, bi, &cnt, k, or_d2(qx, qy, x[j], y[j]), j);
                    }
                }
            }
            int done_all = (ax <= 0 && bx >= ncx - 1 && ay <= 0 && by >= ncy - 1);
            if (done_all) break;
            if (cnt == k) {
                /* distance from q to the nearest not-yet-visited cell */
                double dm = INFINITY;
                if (ax > 0) dm = fmin(dm, qx - (x0 + (double)ax * cs));
                if (bx < ncx - 1) dm = fmin(dm, (x0 + (double)(bx + 1) * cs) - qx);
                if (ay > 0) dm = fmin(dm, qy - (y0 + (double)ay * cs));
                if (by < ncy - 1) dm = fmin(dm, (y0 + (double)(by + 1) * cs) - qy);
                dm *= (1.0 - 1e-9);
                if (dm > 0 && bd[k - 1] < dm * dm) break;
            }
        }
        for (int s = 0; s < k; ++s) {
            idx[i * k + s] = bi[s];
            if (d2) d2[i * k + s] = bd[s];
        }
    }
    free(bd);
    free(bi);
    free(fill);
    free(order);
    free(cell);
    free(start);
    return 0;
}

/* Julia's pairwise reduction (Base.mapreduce_impl, block size 1024),
 * inclusive index range [lo, hi]. */
double or_pairwise_sum(const double* a, int64_t lo, int64_t hi)
{
    if (lo == hi) return a[lo];
    if (hi - lo < 1024) {
        double v = a[lo] + a[lo + 1];
        for (int64_t i = lo + 2; i <= hi; ++i) v = v + a[i];
        return v;
    }
    int64_t mid = lo + ((hi - lo) >> 1);
    return or_pairwise_sum(a, lo, mid) + or_pairwise_sum(a, mid + 1, hi);
}

/* dx_min = minimum(dists_y)[2], dx_avg = mean(dists_y)[2] with dists_y the
 * 2-NN Euclidean distances: geometry_primatives.jl:332-334.  Input is the
 * squared distance to the nearest other node, one per node. */
void or_dx_stats(const double* d2_nn, int64_t N, double* dx_min, double* dx_avg)
{
    double* d = (double*)malloc(sizeof(double) * N);
    double mn = INFINITY;
    for (int64_t i = 0; i < N; ++i) {
        d[i] = sqrt(d2_nn[i]);
        if (d[i] < mn) mn = d[i];
    }
    *dx_min = mn;
    *dx_avg = or_pairwise_sum(d, 0, N - 1) / (double)N;
    free(d);
}

/* ======================================================================
 * Weights.  Reference: compute_flux_operator, compute_operators.jl:409-453
 * (first derivatives) and :549-594 (k-th derivatives).
 * ==================================================================== */

/* Bunch-Kaufman factorization A = U D U^T of a symmetric matrix stored in
 * the upper triangle (column-major, 1-based macros below).  Restates LAPACK
 * DSYTF2 (UPLO='U'), the unblocked path DSYTRF takes for n < 64; Julia's
 * `Symmetric(M) \ rhs` (compute_operators.jl:435) factors with DSYTRF. */
#define AM(i, j) a[((i)-1) + (size_t)((j)-1) * lda]
int or_bk_factor(int n, double* a, int lda, int* ipiv)
{
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    int info = 0;
    int k = n;
    while (k >= 1) {
        int kstep = 1;
        const double absakk = fabs(AM(k, k));
        int imax = 0;
        double colmax = 0.0;
        if (k > 1) {
            imax = 1;
            colmax = fabs(AM(1, k));
            for (int i = 2; i <= k - 1; ++i)
                if (fabs(AM(i, k)) > colmax) {
                    colmax = fabs(AM(i, k));
                    imax = i;
                }
        }
        int kp;
        if (fmax(absakk, colmax) == 0.0 || isnan(absakk)) {
            if (info == 0) info = k;
            kp = k;
        } else {
            if (absakk >= alpha * colmax) {
                kp = k;
            } else {
                int jmax = imax + 1;
                double rowmax = fabs(AM(imax, imax + 1));
                for (int j = imax + 2; j <= k; ++j)
                    if (fabs(AM(imax, j)) > rowmax) {
                        rowmax = fabs(AM(imax, j));
                        jmax = j;
                    }
                if (imax > 1) {
                    jmax = 1;
                    double cm = fabs(AM(1, imax));
                    for (int i = 2; i <= imax - 1; ++i)
                        if (fabs(AM(i, imax)) > cm) {
                            cm = fabs(AM(i, imax));
                            jmax = i;
                        }
                    rowmax = fmax(rowmax, fabs(AM(jmax, imax)));
                }
                if (absakk >= alpha * colmax * (colmax / rowmax)) {
                    kp = k;
                } else if (fabs(AM(imax, imax)) >= alpha * rowmax) {
                    kp = imax;
                } else {
                    kp = imax;
                    kstep = 2;
                }
            }
            const int kk = k - kstep + 1;
            if (kp != kk) {
                for (int i = 1; i <= kp - 1; ++i) {
                    double t = AM(i, kk);
                    AM(i, kk) = AM(i, kp);
                    AM(i, kp) = t;
                }
                for (int j = kp + 1; j <= kk - 1; ++j) {
                    double t = AM(j, kk);
                    AM(j, kk) = AM(kp, j);
                    AM(kp, j) = t;
                }
                double t = AM(kk, kk);
                AM(kk, kk) = AM(kp, kp);
                AM(kp, kp) = t;
                if (kstep == 2) {
                    t = AM(k - 1, k);
                    AM(k - 1, k) = AM(kp, k);
                    AM(kp, k) = t;
                }
            }
            if (kstep == 1) {
                /* DSYR(U, k-1, -r1, A(1,k), 1, A) then DSCAL(k-1, r1, A(1,k)) */
                const double r1 = 1.0 / AM(k, k);
                for (int j = 1; j <= k - 1; ++j) {
                    if (AM(j, k) != 0.0) {
                        const double temp = -r1 * AM(j, k);
                        for (int i = 1; i <= j; ++i) AM(i, j) = AM(i, j) + AM(i, k) * temp;
                    }
                }
                for (int i = 1; i <= k - 1; ++i) AM(i, k) = r1 * AM(i, k);
            } else if (k > 2) {
                double d12 = AM(k - 1, k);
                const double d22 = AM(k - 1, k - 1) / d12;
                const double d11 = AM(k, k) / d12;
                const double t = 1.0 / (d11 * d22 - 1.0);
                d12 = t / d12;
                for (int j = k - 2; j >= 1; --j) {
                    const double wkm1 = d12 * (d11 * AM(j, k - 1) - AM(j, k));
                    const double wk = d12 * (d22 * AM(j, k) - AM(j, k - 1));
                    for (int i = j; i >= 1; --i)
                        AM(i, j) = AM(i, j) - AM(i, k) * wk - AM(i, k - 1) * wkm1;
                    AM(j, k) = wk;
                    AM(j, k - 1) = wkm1;
                }
            }
        }
        if (kstep == 1) {
            ipiv[k - 1] = kp;
        } else {
            ipiv[k - 1] = -kp;
            ipiv[k - 2] = -kp;
        }
        k -= kstep;
    }
    return info;
}

/* Solve with the factorization above: restates LAPACK DSYTRS (UPLO='U'). */
#define BM(i, j) b[((i)-1) + (size_t)((j)-1) * ldb]
void or_bk_solve(int n, const double* a, int lda, const int* ipiv, int nrhs, double* b, int ldb)
{
    int k = n;
    while (k >= 1) {
        if (ipiv[k - 1] > 0) {
            const int kp = ipiv[k - 1];
            if (kp != k)
                for (int j = 1; j <= nrhs; ++j) {
                    double t = BM(k, j);
                    BM(k, j) = BM(kp, j);
                    BM(kp, j) = t;
                }
            for (int j = 1; j <= nrhs; ++j) {
                if (BM(k, j) != 0.0) {
                    const double temp = -1.0 * BM(k, j);
                    for (int i = 1; i <= k - 1; ++i) BM(i, j) = BM(i, j) + AM(i, k) * temp;
                }
            }
            const double r = 1.0 / AM(k, k);
            for (int j = 1; j <= nrhs; ++j) BM(k, j) = r * BM(k, j);
            k -= 1;
        } else {
            const int kp = -ipiv[k - 1];
            if (kp != k - 1)
                for (int j = 1; j <= nrhs; ++j) {
                    double t = BM(k - 1, j);
                    BM(k - 1, j) = BM(kp, j);
                    BM(kp, j) = t;
                }
            for (int j = 1; j <= nrhs; ++j) {
                if (BM(k, j) != 0.0) {
                    const double temp = -1.0 * BM(k, j);
                    for (int i = 1; i <= k - 2; ++i) BM(i, j) = BM(i, j) + AM(i, k) * temp;
                }
            }
            for (int j = 1; j <= nrhs; ++j) {
                if (BM(k - 1, j) != 0.0) {
                    const double temp = -1.0 * BM(k - 1, j);
                    for (int i = 1; i <= k - 2; ++i) BM(i, j) = BM(i, j) + AM(i, k - 1) * temp;
                }
            }
            const double akm1k = AM(k - 1, k);
            const double akm1 = AM(k - 1, k - 1) / akm1k;
            const double ak = AM(k, k) / akm1k;
            const double denom = akm1 * ak - 1.0;
            for (int j = 1; j <= nrhs; ++j) {
                const double bkm1 = BM(k - 1, j) / akm1k;
                const double bk = BM(k, j) / akm1k;
                BM(k - 1, j) = (ak * bkm1 - bk) / denom;
                BM(k, j) = (akm1 * bk - bkm1) / denom;
            }
            k -= 2;
        }
    }
    k = 1;
    while (k <= n) {
        if (ipiv[k - 1] > 0) {
            for (int j = 1; j <= nrhs; ++j) {
                double temp = 0.0;
                for (int i = 1; i <= k - 1; ++i) temp = temp + BM(i, j) * AM(i, k);
                BM(k, j) = BM(k, j) + -1.0 * temp;
            }
            const int kp = ipiv[k - 1];
            if (kp != k)
                for (int j = 1; j <= nrhs; ++j) {
                    double t = BM(k, j);
                    BM(k, j) = BM(kp, j);
                    BM(kp, j) = t;
                }
            k += 1;
        } else {
            for (int j = 1; j <= nrhs; ++j) {
                double temp = 0.0;
                for (int i = 1; i <= k - 1; ++i) temp = temp + BM(i, j) * AM(i, k);
                BM(k, j) = BM(k, j) + -1.0 * temp;
            }
            for (int j = 1; j <= nrhs; ++j) {
                double temp = 0.0;
                for (int i = 1; i <= k - 1; ++i) temp = temp + BM(i, j) * AM(i, k + 1);
                BM(k + 1, j) = BM(k + 1, j) + -1.0 * temp;
            }
            const int kp = -ipiv[k - 1];
            if (kp != k)
                for (int j = 1; j <= nrhs; ++j) {
                    double t = BM(k, j);
                    BM(k, j) = BM(kp, j);
                    BM(kp, j) = t;
                }
            k += 2;
        }
    }
}
#undef AM
#undef BM

/* Number of monomials of total degree <= deg in 2-D: C(deg+2, 2). */
int or_npoly(int deg) { return (deg + 1) * (deg + 2) / 2; }

/* Monomial t of the degree-deg basis: total degree ascending, then x-power
 * descending: 1, x, y, x^2, xy, y^2, x^3, x^2y, xy^2, y^3, ...
 * (the reference's DynamicPolynomials ordering only permutes these). */
static void or_mono_exp(int t, int* ax, int* ay)
{
    int d = 0;
    while (t >= d + 1) {
        t -= d + 1;
        ++d;
    }
    *ax = d - t;
    *ay = t;
}

static inline double or_ipow(double v, int e)
{
    double r = 1.0;
    for (int i = 0; i < e; ++i) r = r * v;
    return r;
}

/* k-th derivative along x (dir 0) or y (dir 1) of phi = r^p, r = |(x,y)|,
 * as sum_t c_t * s^a_t * r^b_t with s the differentiated coordinate:
 * d/ds (s^a r^b) = a s^(a-1) r^b + b s^(a+1) r^(b-2).  The reference gets
 * the same function from Symbolics (compute_operators.jl:17-27, :53-65). */
#define OR_MAXTERM 16
double or_phs_deriv(int p, int k, int dir, double x, double y)
{
    double cc[OR_MAXTERM], c2[OR_MAXTERM];
    int aa[OR_MAXTERM], bb[OR_MAXTERM], a2[OR_MAXTERM], b2[OR_MAXTERM];
    int nt = 1;
    cc[0] = 1.0;
    aa[0] = 0;
    bb[0] = p;
    for (int it = 0; it < k; ++it) {
        int n2 = 0;
        for (int t = 0; t < nt; ++t) {
            if (aa[t] > 0) {
                c2[n2] = cc[t] * aa[t];
                a2[n2] = aa[t] - 1;
                b2[n2] = bb[t];
                ++n2;
            }
            if (bb[t] != 0) {
                c2[n2] = cc[t] * bb[t];
                a2[n2] = aa[t] + 1;
                b2[n2] = bb[t] - 2;
                ++n2;
            }
        }
        /* merge equal (a, b) */
        nt = 0;
        for (int t = 0; t < n2; ++t) {
            int f = -1;
            for (int q = 0; q < nt; ++q)
                if (aa[q] == a2[t] && bb[q] == b2[t]) f = q;
            if (f >= 0) {
                cc[f] += c2[t];
            } else {
                cc[nt] = c2[t];
                aa[nt] = a2[t];
                bb[nt] = b2[t];
                ++nt;
            }
        }
    }
    const double s = dir == 0 ? x : y;
    const double r = sqrt(x * x + y * y);
    double val = 0.0;
    for (int t = 0; t < nt; ++t) {
        if (cc[t] == 0.0) continue;
        double term = cc[t] * or_ipow(s, aa[t]);
        if (bb[t] >= 0)
            term = term * or_ipow(r, bb[t]);
        else
            term = term / or_ipow(r, -bb[t]);
        val = val + term;
    }
    return val;
}

/* s^k nearly correctly rounded, as Julia's Float64^Int power (pow_body). */
static double or_pow_k(double s, int k)
{
    long double r = 1.0L;
    for (int i = 0; i < k; ++i) r *= (long double)s;
    return (double)r;
}

/* Weights for every stencil (ragged: stencil e = idx[ptr[e] .. ptr[e+1]),
 * first entry = centre).  wx/wy are stored like idx.  Per stencil:
 * shift_stencil (:225-246), rbf_block (:191-203), poly_block (:211-223),
 * interpolation_block (:265-267), poly_linearoperator (:283-291),
 * mirror_stencil (:248-263), rbf_linearoperator (:345-358),
 * assemble_rhs (:381-384), M \ rhs (:435), scaling (:438-439 / :579-580).
 * The interpolation column E (:440) is discarded by the reference and not
 * formed here.  Returns 0, or 1 + e for the first singular stencil. */
int or_weights(const double* x, const double* y, int64_t N, const int64_t* ptr,
               const int32_t* idx, int phs_p, int poly_deg, int deriv_k, double* wx,
               double* wy)
{
    const int np = or_npoly(poly_deg);
    int nmax = 0;
    for (int64_t e = 0; e < N; ++e) {
        int n = (int)(ptr[e + 1] - ptr[e]);
        if (n > nmax) nmax = n;
    }
    const int mmax = nmax + np;
    double* M = (double*)malloc(sizeof(double) * mmax * mmax);
    double* B = (double*)malloc(sizeof(double) * mmax * 2);
    double* X = (double*)malloc(sizeof(double) * nmax * 2);
    int* ipiv = (int*)malloc(sizeof(int) * mmax);
    int kfact = 1;
    for (int i = 2; i <= deriv_k; ++i) kfact *= i;
    int rc = 0;
    for (int64_t e = 0; e < N; ++e) {
        const int64_t off = ptr[e];
        const int n = (int)(ptr[e + 1] - off);
        const int m = n + np;
        const int32_t c0 = idx[off];
        double mx = 0.0, my = 0.0;
        for (int j = 0; j < n; ++j) {
            const int32_t q = idx[off + j];
            X[2 * j] = x[q] - x[c0];
            X[2 * j + 1] = y[q] - y[c0];
        }
        for (int j = 0; j < n; ++j) {
            if (fabs(X[2 * j]) > mx) mx = fabs(X[2 * j]);
            if (fabs(X[2 * j + 1]) > my) my = fabs(X[2 * j + 1]);
        }
        const double sx = 1.0 / mx, sy = 1.0 / my;
        for (int j = 0; j < n; ++j) {
            X[2 * j] = X[2 * j] * sx;
            X[2 * j + 1] = X[2 * j + 1] * sy;
        }
        memset(M, 0, sizeof(double) * m * m);
        for (int jj = 0; jj < n; ++jj)
            for (int ii = 0; ii < n; ++ii) {
                const double dx = X[2 * ii] - X[2 * jj], dy = X[2 * ii + 1] - X[2 * jj + 1];
                M[ii + (size_t)jj * m] = or_ipow(sqrt(dx * dx + dy * dy), phs_p);
            }
        for (int t = 0; t < np; ++t) {
            int ax, ay;
            or_mono_exp(t, &ax, &ay);
            for (int ii = 0; ii < n; ++ii) {
                const double v = or_ipow(X[2 * ii], ax) * or_ipow(X[2 * ii + 1], ay);
                M[ii + (size_t)(n + t) * m] = v;
                M[(n + t) + (size_t)ii * m] = v;
            }
        }
        /* right-hand sides: L phi(x_c - x_j) at the mirrored points, the
         * centre replaced by (eps, eps); L monomial at the (shifted) centre */
        for (int j = 0; j < n; ++j) {
            double px = -X[2 * j], py = -X[2 * j + 1];
            if (j == 0) px = py = DBL_EPSILON;
            B[j] = or_phs_deriv(phs_p, deriv_k, 0, px, py);
            B[j + m] = or_phs_deriv(phs_p, deriv_k, 1, px, py);
        }
        for (int t = 0; t < np; ++t) {
            int ax, ay;
            or_mono_exp(t, &ax, &ay);
            B[n + t] = (ax == deriv_k && ay == 0) ? (double)kfact : 0.0;
            B[n + t + m] = (ay == deriv_k && ax == 0) ? (double)kfact : 0.0;
        }
        if (or_bk_factor(m, M, m, ipiv) != 0 && rc == 0) rc = (int)(1 + e);
        or_bk_solve(m, M, m, ipiv, 2, B, m);
        const double fx = or_pow_k(sx, deriv_k), fy = or_pow_k(sy, deriv_k);
        for (int j = 0; j < n; ++j) {
            wx[off + j] = fx * B[j];
            wy[off + j] = fy * B[j + m];
        }
    }
    free(M);
    free(B);
    free(X);
    free(ipiv);
    return rc;
}

/* or_weights over chunks of 64 stencils on OpenMP threads: each stencil is
 * computed by the same code as in or_weights, so the output is the same
 * bits.  Returns 0, or 1 + the first singular stencil's index. */
int or_weights_par(const double* x, const double* y, int64_t N, const int64_t* ptr,
                   const int32_t* idx, int phs_p, int poly_deg, int deriv_k, double* wx,
                   double* wy, int nthreads)
{
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    const int64_t nch = (N + 63) / 64;
    int64_t first = 0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int64_t ch = 0; ch < nch; ++ch) {
        const int64_t e0 = ch * 64, n = (N - e0 < 64) ? N - e0 : 64;
        const int r = or_weights(x, y, n, ptr + e0, idx, phs_p, poly_deg, deriv_k, wx, wy);
        if (r) {
#pragma omp critical
            if (first == 0 || e0 + r < first) first = e0 + r;
        }
    }
    return (int)first;
}

/* ======================================================================
 * Sparse operators.  sparse(rows, cols, vals) (compute_operators.jl:443-452)
 * gives CSC sorted by column then row; the products below restate
 * SparseArrays' mul!(C, A, x, alpha, beta) and mul!(C, A', x, alpha, beta)
 * loops (pointcloudsolver/rbfsolver.jl:21-27 calls them).
 * ==================================================================== */
/* CSC pattern of the matrix with rows e and columns idx[ptr[e]..].  perm[q]
 * is the ELL/ragged position of CSC entry q. */
void or_csc_build(int64_t N, const int64_t* ptr, const int32_t* idx, int64_t* colptr,
                  int32_t* rowval, int64_t* perm)
{
    memset(colptr, 0, sizeof(int64_t) * (N + 1));
    for (int64_t e = 0; e < N; ++e)
        for (int64_t q = ptr[e]; q < ptr[e + 1]; ++q) colptr[idx[q] + 1]++;
    for (int64_t c = 0; c < N; ++c) colptr[c + 1] += colptr[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * N);
    memcpy(fill, colptr, sizeof(int64_t) * N);
    /* rows visited in ascending order => each column's rows come out sorted */
    for (int64_t e = 0; e < N; ++e)
        for (int64_t q = ptr[e]; q < ptr[e + 1]; ++q) {
            int64_t pos = fill[idx[q]]++;
            rowval[pos] = (int32_t)e;
            perm[pos] = q;
        }
    free(fill);
}

/* out = (overwrite ? 0 : out) + A * (x * alpha), column by column */
void or_csc_mul(int64_t N, const int64_t* colptr, const int32_t* rowval, const double* nzval,
                const double* x, double* out, double alpha, int overwrite)
{
    if (overwrite) memset(out, 0, sizeof(double) * N);
    for (int64_t col = 0; col < N; ++col) {
        const double axj = x[col] * alpha;
        for (int64_t j = colptr[col]; j < colptr[col + 1]; ++j)
            out[rowval[j]] = out[rowval[j]] + nzval[j] * axj;
    }
}

/* out += (A^T x) * alpha, one dot product per column of A */
void or_csc_tmul(int64_t N, const int64_t* colptr, const int32_t* rowval,
                 const double* nzval, const double* x, double* out, double alpha)
{
    for (int64_t col = 0; col < N; ++col) {
        double tmp = 0.0;
        for (int64_t j = colptr[col]; j < colptr[col + 1]; ++j)
            tmp = tmp + nzval[j] * x[rowval[j]];
        out[col] = out[col] + tmp * alpha;
    }
}

/* ======================================================================
 * rhs!  Reference: Trixi.rhs!, pointcloudsolver/rbfsolver.jl:397-428.
 * State is SoA: component v of node e at u[v*N + e] (StructArray layout,
 * rbfsolver.jl:111-116).
 * ==================================================================== */
/* Physical flux.  Restates Trixi's flux(u, orientation, equations) for
 * CompressibleEulerEquations2D (evaluated inside Trixi's @muladd block, so
 * its multiply-adds are fused) and LinearScalarAdvectionEquation2D; called
 * from calc_fluxes!, rbfsolver.jl:258-260. */
void or_flux(int eq, const double* adv_a, double gamma, int orientation, const double* u,
             double* f)
{
    if (eq == OR_EQ_ADVECTION) {
        f[0] = adv_a[orientation] * u[0];
        return;
    }
    const double rho = u[0], m1 = u[1], m2 = u[2], E = u[3];
    const double v1 = m1 / rho, v2 = m2 / rho;
    const double p = (gamma - 1.0) * fma(-0.5, fma(m1, v1, m2 * v2), E);
    if (orientation == 0) {
        f[0] = m1;
        f[1] = fma(m1, v1, p);
        f[2] = m1 * v2;
        f[3] = (E + p) * v1;
    } else {
        f[0] = m2;
        f[1] = m2 * v1;
        f[2] = fma(m2, v2, p);
        f[3] = (E + p) * v2;
    }
}

/* Boundary/initial value functions, evaluated at (x, y, t).
 *   CONST:    params[0..nvar)
 *   PULSE:    scalar exp(-((x - a1 t - x0)^2 + (y - a2 t - y0)^2) / w^2),
 *             params (x0, y0, w, a1, a2)
 *   WAVE:     Euler density wave rho = r0 + A sin(2 pi (x + y - t (v1 + v2)) + phi),
 *             velocity (v1, v2), pressure p; params (r0, A, v1, v2, p, phi)
 *   GRADIENT: Euler linear state of the reference's operator tests
 *             (test/divergence_test.jl:26-36) */
void or_bc_value(int fn, const double* q, int nvar, double gamma, double x, double y, double t,
                 double* out)
{
    if (fn == OR_FN_CONST) {
        for (int v = 0; v < nvar; ++v) out[v] = q[v];
    } else if (fn == OR_FN_PULSE) {
        const double dx = (x - q[3] * t) - q[0], dy = (y - q[4] * t) - q[1];
        out[0] = exp(-(dx * dx + dy * dy) / (q[2] * q[2]));
    } else if (fn == OR_FN_WAVE) {
        const double arg = (x + y) - t * (q[2] + q[3]);
        const double rho = q[0] + q[1] * sin(2.0 * M_PI * arg + q[5]);
        out[0] = rho;
        out[1] = rho * q[2];
        out[2] = rho * q[3];
        out[3] = q[4] / (gamma - 1.0) + 0.5 * rho * (q[2] * q[2] + q[3] * q[3]);
    } else if (fn == OR_FN_GRADIENT) {
        out[0] = 1.4 + 0.1 * x + 0.1 * y;
        out[1] = 4.1 + 0.1 * x + 0.1 * y;
        out[2] = 0.0 + 0.1 * x + 0.1 * y;
        out[3] = 8.8 + 0.1 * x + 0.1 * y;
    }
}

/* One pass of calc_boundary_flux! over all tags (rbfsolver.jl:277-318),
 * with the strongly imposed conditions of PointCloudBCs.jl:
 *   Dirichlet (:49-63): u_b <- g(x_b, t), du_b <- 0
 *   slip wall (:87-106, :15-21): momentum projected on the wall,
 *             du_b momentum <- 0
 *   do-nothing (:108-115): unchanged */
void or_bc_pass(const or_rhs_cfg* c, double t, double* u, double* du)
{
    const int64_t N = c->N;
    for (int g = 0; g < c->ntags; ++g) {
        const int kind = c->tag_kind[g];
        for (int32_t q = c->tag_ptr[g]; q < c->tag_ptr[g + 1]; ++q) {
            const int32_t b = c->bidx[q];
            if (kind == OR_BC_DIRICHLET) {
                double val[4];
                if (c->tag_fn[g] == OR_FN_TABLE) {
                    for (int v = 0; v < c->nvar; ++v) val[v] = c->btable[(int64_t)q * c->nvar + v];
                } else {
                    or_bc_value(c->tag_fn[g], c->tag_fn_params + 8 * g, c->nvar, c->gamma,
                                c->px[b], c->py[b], t, val);
                }
                for (int v = 0; v < c->nvar; ++v) {
                    u[v * N + b] = val[v];
                    du[v * N + b] = 0.0;
                }
            } else if (kind == OR_BC_SLIP) {
                const double n1 = c->bnormal[2 * q], n2 = c->bnormal[2 * q + 1];
                const double nrm = sqrt(n1 * n1 + n2 * n2);
                const double e1 = n1 / nrm, e2 = n2 / nrm;
                const double m1 = u[1 * N + b], m2 = u[2 * N + b];
                const double s = m1 * e1 + m2 * e2;
                u[1 * N + b] = fma(-s, e1, m1);
                u[2 * N + b] = fma(-s, e2, m2);
                du[1 * N + b] = 0.0;
                du[2 * N + b] = 0.0;
            }
        }
    }
}

/* update_upwind_visc!, hyperviscosity.jl:246-285 (no @muladd there) */
void or_upwind_visc(const or_rhs_cfg* c, const double* u, double* eps_uw)
{
    const int64_t N = c->N;
    const double gam = c->gamma;
    for (int64_t i = 0; i < N; ++i) {
        const double rho = u[i], m1 = u[N + i], m2 = u[2 * N + i], E = u[3 * N + i];
        /* cons2prim (Trixi, @muladd) */
        const double v1 = m1 / rho, v2 = m2 / rho;
        const double p = (gam - 1.0) * fma(-0.5, fma(m1, v1, m2 * v2), E);
        const double speed = sqrt(v1 * v1 + v2 * v2);
        double sound = 0.0;
        if (!(p < 0.0 || rho < 0.0)) sound = sqrt(gam * p / rho);
        eps_uw[i] = c->c_uw * 0.5 * c->dx_avg * (speed + sound);
    }
}

/* ode_mean (auxiliary/mpi.jl:40-52: sum over the StructArray divided by
 * recursive_length = nvar*N) and ode_maximum (mpi.jl:71-82: `maximum` of
 * SVectors, i.e. the lexicographic maximum) of |u_i - mean|, with exact
 * zeros replaced by eps() (hyperviscosity.jl:305-311). */
void or_rv_norms(int64_t N, int nvar, const double* u, double* mean, double* ninf)
{
    for (int v = 0; v < nvar; ++v)
        mean[v] = or_pairwise_sum(u + v * N, 0, N - 1) / (double)(nvar * N);
    double best[4];
    for (int64_t i = 0; i < N; ++i) {
        double cur[4];
        for (int v = 0; v < nvar; ++v) cur[v] = fabs(u[v * N + i] - mean[v]);
        int take = (i == 0);
        if (!take) {
            /* max(x, y) = isless(y, x) ? x : y, lexicographic isless */
            for (int v = 0; v < nvar; ++v) {
                if (best[v] < cur[v]) {
                    take = 1;
                    break;
                }
                if (best[v] > cur[v]) break;
                if (v == nvar - 1) take = 1; /* equal: keep the later one */
            }
        }
        if (take)
            for (int v = 0; v < nvar; ++v) best[v] = cur[v];
    }
    for (int v = 0; v < nvar; ++v) ninf[v] = (best[v] == 0.0) ? DBL_EPSILON : best[v];
}

/* Julia's max(x, y) for floats: NaN wins, otherwise the larger value */
static inline double or_jmax(double x, double y)
{
    if (isnan(x)) return x;
    if (isnan(y)) return y;
    return (y > x || (signbit(y) < signbit(x))) ? y : x;
}

/* update_residual_visc!, hyperviscosity.jl:289-329 */
void or_residual_visc(const or_rhs_cfg* c, const double* du, const double* u, double* eps_rv)
{
    const int64_t N = c->N;
    double mean[4], ninf[4];
    or_rv_norms(N, 4, u, mean, ninf);
    const double h = c->dx_avg;
    for (int64_t i = 0; i < N; ++i) {
        double mr = 0.0;
        for (int v = 0; v < 4; ++v) {
            const double res = fabs(c->approx_du[v * N + i] - du[v * N + i]);
            const double r = res / ninf[v];
            mr = (v == 0) ? r : or_jmax(mr, r);
        }
        eps_rv[i] = 0.5 * c->c_rv * (h * h) * mr;
    }
}

/* update_visc!, hyperviscosity.jl:331-349 */
void or_update_visc(int64_t N, const double* eps_uw, const double* eps_rv, int success_iter,
                    double* eps, int8_t* eps_c)
{
    for (int64_t i = 0; i < N; ++i) {
        if (isnan(eps_rv[i]) || isinf(eps_rv[i]) || success_iter == 0) {
            if (isnan(eps_uw[i]) || isinf(eps_uw[i])) {
                eps[i] = DBL_EPSILON;
                eps_c[i] = 2;
            } else {
                eps[i] = eps_uw[i];
                eps_c[i] = 1;
            }
        } else {
            eps[i] = (eps_rv[i] < eps_uw[i]) ? eps_rv[i] : eps_uw[i];
            eps_c[i] = eps_rv[i] < eps_uw[i] ? 0 : 1;
        }
    }
}

int64_t or_rhs_scratch_len(const or_rhs_cfg* c)
{
    return (int64_t)c->nvar * c->N + 4 * c->N;
}

/* Flyer hyperviscosity, hyperviscosity.jl:52-64: du += (H u) * (-gamma),
 * H = Dx^2k + Dy^2k, applied like mul_by_accum! (one CSC product). */
static void or_src_flyer(const or_rhs_cfg* c, const double* u, double* du)
{
    for (int v = 0; v < c->nvar; ++v)
        or_csc_mul(c->N, c->colptr, c->rowval, c->hv, u + v * c->N, du + v * c->N, -c->hv_gamma, 0);
}

/* Tominec hyperviscosity, hyperviscosity.jl:80-134: the reference forms
 * H = lap' * lap (lap = Dxx + Dyy) explicitly as a sparse product and
 * applies it the same way as Flyer; here H u = lap' (lap u): local = lap u,
 * du += (lap' local) * (-gamma).  The two forms agree to rounding; the
 * oracle tests pin this one against scipy's explicit lap.T @ lap. */
static void or_src_tominec(const or_rhs_cfg* c, const double* u, double* du, double* local)
{
    for (int v = 0; v < c->nvar; ++v) {
        or_csc_mul(c->N, c->colptr, c->rowval, c->lap, u + v * c->N, local, 1.0, 1);
        or_csc_tmul(c->N, c->colptr, c->rowval, c->lap, local, du + v * c->N, -c->tom_gamma);
    }
}

/* ======================================================================
 * IGR source, sources/IGR.jl.  Euler equations; state SoA.
 * ==================================================================== */
/* update_igr_rhs! (IGR.jl:117-157): primitive variables (Trixi cons2prim:
 * v = rho_v / rho), their derivatives D_d prim by mul_by! (one CSC product
 * per field, column order), then per node
 *   trace    = flux_x[1] + flux_y[2]           (D_x rho + D_y v1, as written)
 *   trace_sq = flux_x[2]^2 + 2 flux_y[2] flux_x[3] + flux_y[3]^2
 *   b        = alpha (trace * trace + trace_sq)
 * and rho_inv = 1 / rho (update_sigma!, :176-178). */
void or_igr_rhs(const or_rhs_cfg* c, const double* u, double* b, double* rinv)
{
    const int64_t N = c->N;
    double* v1 = (double*)malloc(sizeof(double) * N);
    double* v2 = (double*)malloc(sizeof(double) * N);
    double* f = (double*)calloc((size_t)5 * N, sizeof(double));
    for (int64_t i = 0; i < N; ++i) {
        v1[i] = u[N + i] / u[i];
        v2[i] = u[2 * N + i] / u[i];
    }
    double* dxr = f;
    double* dxv1 = f + N;
    double* dxv2 = f + 2 * N;
    double* dyv1 = f + 3 * N;
    double* dyv2 = f + 4 * N;
    or_csc_mul(N, c->colptr, c->rowval, c->dx, u, dxr, 1.0, 1);
    or_csc_mul(N, c->colptr, c->rowval, c->dx, v1, dxv1, 1.0, 1);
    or_csc_mul(N, c->colptr, c->rowval, c->dx, v2, dxv2, 1.0, 1);
    or_csc_mul(N, c->colptr, c->rowval, c->dy, v1, dyv1, 1.0, 1);
    or_csc_mul(N, c->colptr, c->rowval, c->dy, v2, dyv2, 1.0, 1);
    for (int64_t i = 0; i < N; ++i) {
        double trace = 0.0;
        trace += dxr[i];
        trace += dyv1[i];
        const double tsq = dxv1[i] * dxv1[i] + 2.0 * dyv1[i] * dxv2[i] + dyv2[i] * dyv2[i];
        b[i] = c->igr_alpha * (trace * trace + tsq);
        rinv[i] = 1.0 / u[i];
    }
    free(v1);
    free(v2);
    free(f);
}

/* IGRCompositeMatrix mul! (IGR.jl:50-66):
 *   tmp1 = Dx v; tmp2 = Dy v; tmp1 = rinv .* tmp1; tmp2 = rinv .* tmp2;
 *   tmp3 = Dx tmp1; tmp1 = Dy tmp2; y = rinv .* v; y -= alpha .* (tmp3 + tmp1) */
void or_igr_apply(const or_rhs_cfg* c, const double* rinv, const double* v, double* y)
{
    const int64_t N = c->N;
    double* t1 = (double*)malloc(sizeof(double) * N);
    double* t2 = (double*)malloc(sizeof(double) * N);
    double* t3 = (double*)malloc(sizeof(double) * N);
    or_csc_mul(N, c->colptr, c->rowval, c->dx, v, t1, 1.0, 1);
    or_csc_mul(N, c->colptr, c->rowval, c->dy, v, t2, 1.0, 1);
    for (int64_t i = 0; i < N; ++i) {
        t1[i] = rinv[i] * t1[i];
        t2[i] = rinv[i] * t2[i];
    }
    or_csc_mul(N, c->colptr, c->rowval, c->dx, t1, t3, 1.0, 1);
    or_csc_mul(N, c->colptr, c->rowval, c->dy, t2, t1, 1.0, 1);
    for (int64_t i = 0; i < N; ++i) {
        y[i] = rinv[i] * v[i];
        y[i] -= c->igr_alpha * (t3[i] + t1[i]);
    }
    free(t1);
    free(t2);
    free(t3);
}

static double or_norm2(int64_t n, const double* r)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += r[i] * r[i];
    return sqrt(s);
}

/* Conjugate gradients as IterativeSolvers.jl's cg! (cg.jl, CGIterable with
 * the identity preconditioner; published algorithm, package version not
 * pinned by the reference's Project.toml): r = b - A x, residual = |r|,
 * tol = max(reltol |r|, 0) with reltol = sqrt(eps); per iteration, unless
 * residual <= tol or maxiter is reached:
 *   beta = residual^2 / prev^2 (prev = 1 at start); u = r + beta u;
 *   c = A u; alpha = residual^2 / dot(u, c); x += alpha u; r -= alpha c;
 *   prev = residual; residual = |r|.
 * The norms and dot products are summed in index order here (BLAS nrm2 and
 * dot order: parity unpinned at rounding level). */
int or_igr_cg(const or_rhs_cfg* c, const double* rinv, const double* b, double* x, int maxiter)
{
    const int64_t N = c->N;
    double* r = (double*)malloc(sizeof(double) * N);
    double* uu = (double*)calloc((size_t)N, sizeof(double));
    double* cc = (double*)malloc(sizeof(double) * N);
    or_igr_apply(c, rinv, x, cc);
    for (int64_t i = 0; i < N; ++i) r[i] = b[i] - cc[i];
    double residual = or_norm2(N, r), prev = 1.0;
    const double tol = sqrt(DBL_EPSILON) * residual;
    int it = 0;
    while (it < maxiter && !(residual <= tol)) {
        const double beta = residual * residual / (prev * prev);
        for (int64_t i = 0; i < N; ++i) uu[i] = r[i] + beta * uu[i];
        or_igr_apply(c, rinv, uu, cc);
        double d = 0.0;
        for (int64_t i = 0; i < N; ++i) d += uu[i] * cc[i];
        const double a = residual * residual / d;
        for (int64_t i = 0; i < N; ++i) {
            x[i] += a * uu[i];
            r[i] -= a * cc[i];
        }
        prev = residual;
        residual = or_norm2(N, r);
        ++it;
    }
    free(r);
    free(uu);
    free(cc);
    return it;
}

/* The source (IGR.jl:212-239): sigma = 0, solve, then for each dimension
 * du += D_d (flux_igr(sigma) * -1) (mul_by_accum!(D_d, -1)); flux_igr puts
 * sigma in component 2 (x) or 3 (y), the others are zero. */
static void or_src_igr(const or_rhs_cfg* c, const double* u, double* du)
{
    const int64_t N = c->N;
    double* b = (double*)malloc(sizeof(double) * N);
    double* rinv = (double*)malloc(sizeof(double) * N);
    double* sigma = (double*)calloc((size_t)N, sizeof(double));
    or_igr_rhs(c, u, b, rinv);
    or_igr_cg(c, rinv, b, sigma, c->igr_maxiter);
    or_csc_mul(N, c->colptr, c->rowval, c->dx, sigma, du + N, -1.0, 0);
    or_csc_mul(N, c->colptr, c->rowval, c->dy, sigma, du + 2 * N, -1.0, 0);
    if (c->igr_sigma_out) memcpy(c->igr_sigma_out, sigma, sizeof(double) * N);
    free(b);
    free(rinv);
    free(sigma);
}

/* residual / upwind viscosity source, hyperviscosity.jl:351-409 */
static void or_src_rv(const or_rhs_cfg* c, const double* u, double* du, double* eps,
                      int8_t* eps_c, double* scratch)
{
    const int64_t N = c->N;
    double* eps_uw = scratch;
    double* eps_rv = scratch + N;
    double* local = scratch + 2 * N;
    or_upwind_visc(c, u, eps_uw);
    if (c->rv_mode == OR_RV_UPWIND) {
        for (int64_t i = 0; i < N; ++i) {
            eps[i] = eps_uw[i];
            eps_c[i] = 1;
        }
    } else {
        or_residual_visc(c, du, u, eps_rv);
        or_update_visc(N, eps_uw, eps_rv, c->success_iter, eps, eps_c);
    }
    const double* D[2] = {c->dx, c->dy};
    for (int d = 0; d < 2; ++d)
        for (int v = 0; v < c->nvar; ++v) {
            or_csc_mul(N, c->colptr, c->rowval, D[d], u + v * N, local, 1.0, 1);
            for (int64_t i = 0; i < N; ++i) local[i] = eps[i] * local[i];
            or_csc_tmul(N, c->colptr, c->rowval, D[d], local, du + v * N, -1.0);
        }
}

int or_rhs(const or_rhs_cfg* c, double t, double* u, double* du, double* eps_out,
           int8_t* eps_c_out, double* scratch)
{
    const int64_t N = c->N;
    const int nv = c->nvar;
    memset(du, 0, sizeof(double) * nv * N); /* reset_du!, rbfsolver.jl:118-124 */
    or_bc_pass(c, t, u, du);                /* first pass, :405-408 */
    /* calc_fluxes!, :247-265: per dimension, flux at every node, then
     * du.x[v] += -D_d * F.x[v] */
    double* F = scratch;
    const double* D[2] = {c->dx, c->dy};
    for (int d = 0; d < 2; ++d) {
        for (int64_t e = 0; e < N; ++e) {
            double ue[4], fe[4];
            for (int v = 0; v < nv; ++v) ue[v] = u[v * N + e];
            or_flux(c->eq, c->adv_a, c->gamma, d, ue, fe);
            for (int v = 0; v < nv; ++v) F[v * N + e] = fe[v];
        }
        for (int v = 0; v < nv; ++v)
            or_csc_mul(N, c->colptr, c->rowval, D[d], F + v * N, du + v * N, -1.0, 0);
    }
    /* calc_sources!, :388-395, in keyword order */
    double* eps = eps_out;
    int8_t* eps_c = eps_c_out;
    double* rv_scr = scratch + (int64_t)nv * N;
    int8_t* own_c = NULL;
    double* own_e = NULL;
    if (c->rv_mode != OR_RV_NONE && (!eps || !eps_c)) {
        own_e = (double*)malloc(sizeof(double) * N);
        own_c = (int8_t*)malloc(N);
        if (!eps) eps = own_e;
        if (!eps_c) eps_c = own_c;
    }
    int order[4] = {0, 0, 0, 0};
    if (c->src_order[0]) {
        for (int q = 0; q < 4; ++q) order[q] = c->src_order[q];
    } else {
        order[0] = c->hv_first ? OR_SRC_FLYER : OR_SRC_VISCOSITY;
        order[1] = c->hv_first ? OR_SRC_VISCOSITY : OR_SRC_FLYER;
        order[2] = OR_SRC_TOMINEC;
        order[3] = OR_SRC_IGR;
    }
    for (int q = 0; q < 4 && order[q]; ++q) {
        if (order[q] == OR_SRC_IGR && c->igr_alpha != 0.0 && c->eq == OR_EQ_EULER) or_src_igr(c, u, du);
        if (order[q] == OR_SRC_FLYER && c->hv) or_src_flyer(c, u, du);
        if (order[q] == OR_SRC_TOMINEC && c->lap) or_src_tominec(c, u, du, rv_scr + 3 * N);
        if (order[q] == OR_SRC_VISCOSITY && c->rv_mode != OR_RV_NONE) or_src_rv(c, u, du, eps, eps_c, rv_scr);
    }
    or_bc_pass(c, t, u, du); /* second pass, :422-425 */
    free(own_e);
    free(own_c);
    return 0;
}

/* ======================================================================
 * Multi-threaded rhs! for the CPU baseline timing only (not a parity
 * reference): the steps of or_rhs (pointcloudsolver/rbfsolver.jl:397-428;
 * sources/hyperviscosity.jl:246-349 for eps, :382-409 for its apply) for
 * the Euler equations with the residual or upwind viscosity and strong
 * boundary conditions, with the forward products as
 * row gathers over the row-form stencil (rptr, ridx, rwx, rwy: the ELL or
 * ragged input of the weights) and the transposed products as column dots
 * over the CSC, each split over OpenMP threads.  Row sums and the mean run
 * in a different order from or_rhs, so the two agree to rounding, not bit
 * for bit.  Returns -1 for configurations it does not cover.  scratch:
 * 17 N doubles.
 * ==================================================================== */
static inline int or_lex_less(const double* a, const double* b)
{
    for (int v = 0; v < 4; ++v) {
        if (a[v] < b[v]) return 1;
        if (a[v] > b[v]) return 0;
    }
    return 0;
}

int or_rhs_par(const or_rhs_cfg* c, const int64_t* rptr, const int32_t* ridx, const double* rwx,
               const double* rwy, double t, double* u, double* du, double* scratch, int nthreads)
{
    if (c->eq != OR_EQ_EULER || c->nvar != 4 || c->hv || c->lap || c->igr_alpha != 0.0) return -1;
    const int64_t N = c->N;
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    memset(du, 0, sizeof(double) * 4 * N);
    or_bc_pass(c, t, u, du);
    double* F = scratch;         /* F1 in 4 N, then F2 in 4 N */
    double* eps = scratch + 8 * N;
    double* G = scratch + 9 * N; /* eps Dx u in 4 N, then eps Dy u in 4 N */
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t e = 0; e < N; ++e) {
        double ue[4], f1[4], f2[4];
        for (int v = 0; v < 4; ++v) ue[v] = u[v * N + e];
        or_flux(c->eq, c->adv_a, c->gamma, 0, ue, f1);
        or_flux(c->eq, c->adv_a, c->gamma, 1, ue, f2);
        for (int v = 0; v < 4; ++v) {
            F[v * N + e] = f1[v];
            F[(4 + v) * N + e] = f2[v];
        }
    }
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t e = 0; e < N; ++e)
        for (int v = 0; v < 4; ++v) {
            double a = 0.0;
            for (int64_t q = rptr[e]; q < rptr[e + 1]; ++q) a = a + rwx[q] * (F[v * N + ridx[q]] * -1.0);
            for (int64_t q = rptr[e]; q < rptr[e + 1]; ++q) a = a + rwy[q] * (F[(4 + v) * N + ridx[q]] * -1.0);
            du[v * N + e] = a;
        }
    if (c->rv_mode != OR_RV_NONE) {
        const double gam = c->gamma, h = c->dx_avg;
        double mean[4] = {0.0, 0.0, 0.0, 0.0}, ninf[4] = {1.0, 1.0, 1.0, 1.0};
        if (c->rv_mode == OR_RV_RESIDUAL) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma omp parallel for schedule(static) num_threads(nt) reduction(+ : s0, s1, s2, s3)
            for (int64_t e = 0; e < N; ++e) {
                s0 += u[e];
                s1 += u[N + e];
                s2 += u[2 * N + e];
                s3 += u[3 * N + e];
            }
            mean[0] = s0 / (double)(4 * N);
            mean[1] = s1 / (double)(4 * N);
            mean[2] = s2 / (double)(4 * N);
            mean[3] = s3 / (double)(4 * N);
            double best[4] = {-1.0, -1.0, -1.0, -1.0};
#pragma omp parallel num_threads(nt)
            {
                double mine[4] = {-1.0, -1.0, -1.0, -1.0};
#pragma omp for schedule(static)
                for (int64_t e = 0; e < N; ++e) {
                    double cur[4];
                    for (int v = 0; v < 4; ++v) cur[v] = fabs(u[v * N + e] - mean[v]);
                    if (or_lex_less(mine, cur))
                        for (int v = 0; v < 4; ++v) mine[v] = cur[v];
                }
#pragma omp critical
                if (or_lex_less(best, mine))
                    for (int v = 0; v < 4; ++v) best[v] = mine[v];
            }
            for (int v = 0; v < 4; ++v) ninf[v] = (best[v] == 0.0) ? DBL_EPSILON : best[v];
        }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t e = 0; e < N; ++e) {
            /* update_upwind_visc! (as or_upwind_visc) */
            const double rho = u[e], m1 = u[N + e], m2 = u[2 * N + e], E = u[3 * N + e];
            const double v1 = m1 / rho, v2 = m2 / rho;
            const double p = (gam - 1.0) * fma(-0.5, fma(m1, v1, m2 * v2), E);
            const double speed = sqrt(v1 * v1 + v2 * v2);
            double sound = 0.0;
            if (!(p < 0.0 || rho < 0.0)) sound = sqrt(gam * p / rho);
            const double eps_uw = c->c_uw * 0.5 * h * (speed + sound);
            if (c->rv_mode == OR_RV_UPWIND) {
                eps[e] = eps_uw;
                continue;
            }
            /* update_residual_visc! and update_visc! (as or_residual_visc, or_update_visc) */
            double mr = 0.0;
            for (int v = 0; v < 4; ++v) {
                const double r = fabs(c->approx_du[v * N + e] - du[v * N + e]) / ninf[v];
                mr = (v == 0) ? r : or_jmax(mr, r);
            }
            const double eps_rv = 0.5 * c->c_rv * (h * h) * mr;
            if (isnan(eps_rv) || isinf(eps_rv) || c->success_iter == 0)
                eps[e] = (isnan(eps_uw) || isinf(eps_uw)) ? DBL_EPSILON : eps_uw;
            else
                eps[e] = (eps_rv < eps_uw) ? eps_rv : eps_uw;
        }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t e = 0; e < N; ++e)
            for (int v = 0; v < 4; ++v) {
                double gx = 0.0, gy = 0.0;
                for (int64_t q = rptr[e]; q < rptr[e + 1]; ++q) {
                    gx = gx + rwx[q] * u[v * N + ridx[q]];
                    gy = gy + rwy[q] * u[v * N + ridx[q]];
                }
                G[v * N + e] = eps[e] * gx;
                G[(4 + v) * N + e] = eps[e] * gy;
            }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t col = 0; col < N; ++col)
            for (int v = 0; v < 4; ++v) {
                double tx = 0.0, ty = 0.0;
                for (int64_t j = c->colptr[col]; j < c->colptr[col + 1]; ++j) {
                    tx = tx + c->dx[j] * G[v * N + c->rowval[j]];
                    ty = ty + c->dy[j] * G[(4 + v) * N + c->rowval[j]];
                }
                du[v * N + col] = (du[v * N + col] + tx * -1.0) + ty * -1.0;
            }
    }
    or_bc_pass(c, t, u, du);
    return 0;
}

/* ======================================================================
 * HistoryCallback, callbacks_step/history.jl:91-152.
 * time_hist[0] is the most recent time; sol_hist holds approx_order+1
 * snapshots of nvals values each, slot 0 the most recent.
 * ==================================================================== */
/* time_deriv_weights!, history.jl:131-152: w = scale * (A' \ b) with the
 * Vandermonde A[:, k] = t_.^(k-1) at t_ = t*scale, b[k] = (k-1) t_1^(k-2).
 * The dense solve restates LAPACK DGETRF/DGETRS (partial pivoting). */
void or_time_deriv_weights(int q, const double* t, double* w)
{
    double A[8 * 8], b[8];
    double mx = 0.0;
    for (int i = 0; i < q; ++i)
        if (fabs(t[i]) > mx) mx = fabs(t[i]);
    const double scale = 1.0 / mx;
    double ts[8];
    for (int i = 0; i < q; ++i) ts[i] = t[i] * scale;
    const double te = ts[0];
    /* At = A' : At[k][i] = ts[i]^k */
    for (int k = 0; k < q; ++k) {
        for (int i = 0; i < q; ++i) A[k + i * q] = (k == 0) ? 1.0 : or_pow_k(ts[i], k);
        b[k] = (k == 0) ? 0.0 : (double)k * ((k == 1) ? 1.0 : or_pow_k(te, k - 1));
    }
    int piv[8];
    for (int j = 0; j < q; ++j) {
        int p = j;
        double pm = fabs(A[j + j * q]);
        for (int i = j + 1; i < q; ++i)
            if (fabs(A[i + j * q]) > pm) {
                pm = fabs(A[i + j * q]);
                p = i;
            }
        piv[j] = p;
        if (p != j)
            for (int c2 = 0; c2 < q; ++c2) {
                double tmp = A[j + c2 * q];
                A[j + c2 * q] = A[p + c2 * q];
                A[p + c2 * q] = tmp;
            }
        const double r = 1.0 / A[j + j * q];
        for (int i = j + 1; i < q; ++i) A[i + j * q] = A[i + j * q] * r;
        for (int c2 = j + 1; c2 < q; ++c2)
            for (int i = j + 1; i < q; ++i)
                A[i + c2 * q] = A[i + c2 * q] - A[i + j * q] * A[j + c2 * q];
    }
    for (int j = 0; j < q; ++j)
        if (piv[j] != j) {
            double tmp = b[j];
            b[j] = b[piv[j]];
            b[piv[j]] = tmp;
        }
    for (int j = 0; j < q; ++j)
        for (int i = j + 1; i < q; ++i) b[i] = b[i] - A[i + j * q] * b[j];
    for (int j = q - 1; j >= 0; --j) {
        b[j] = b[j] / A[j + j * q];
        for (int i = 0; i < j; ++i) b[i] = b[i] - A[i + j * q] * b[j];
    }
    for (int i = 0; i < q; ++i) w[i] = scale * b[i];
}

/* modify_cache!(::SourceResidualViscosityTominec), history.jl:91-129:
 * shift_soln_history! then update_approx_du! (the accumulation sits in
 * the @muladd block, hence fma). */
void or_history_update(int64_t nvals, int approx_order, int success_iter, double t,
                       const double* u, double* time_hist, double* sol_hist, double* approx_du)
{
    const int P = approx_order + 1;
    for (int i = P - 1; i >= 1; --i) {
        time_hist[i] = time_hist[i - 1];
        memcpy(sol_hist + (int64_t)i * nvals, sol_hist + (int64_t)(i - 1) * nvals,
               sizeof(double) * nvals);
    }
    time_hist[0] = t;
    memcpy(sol_hist, u, sizeof(double) * nvals);
    memset(approx_du, 0, sizeof(double) * nvals);
    const int q = (success_iter + 1 < P) ? success_iter + 1 : P;
    if (success_iter > 0) {
        double w[8];
        or_time_deriv_weights(q, time_hist, w);
        for (int i = 0; i < q; ++i)
            for (int64_t k = 0; k < nvals; ++k)
                approx_du[k] = fma(w[i], sol_hist[(int64_t)i * nvals + k], approx_du[k]);
    }
}

/* ======================================================================
 * Zhang-Shu limiter, callbacks_stage/positivity_zhang_shu_point2d.jl:22-83
 * (inside @muladd).  scratch: 2*nvar*N doubles.
 * ==================================================================== */
static double or_variable(int variable, double gamma, const double* s)
{
    if (variable == OR_VAR_DENSITY) return s[0];
    /* Trixi pressure(u, ::CompressibleEulerEquations2D) */
    return (gamma - 1.0) * (s[3] - 0.5 * (s[1] * s[1] + s[2] * s[2]) / s[0]);
}

void or_limiter_zhang_shu(int64_t N, int nvar, double gamma, const int64_t* ptr,
                          const int32_t* idx, double threshold, int variable, double* u,
                          double* scratch)
{
    double* loc = scratch;
    double* mean = scratch + (int64_t)nvar * N;
    memset(scratch, 0, sizeof(double) * 2 * nvar * N);
    for (int64_t e = 0; e < N; ++e) {
        double vmin = INFINITY;
        for (int64_t q = ptr[e]; q < ptr[e + 1]; ++q) {
            double s[4];
            for (int v = 0; v < nvar; ++v) s[v] = u[v * N + idx[q]];
            const double val = or_variable(variable, gamma, s);
            vmin = (val < vmin) ? val : vmin;
        }
        if (!(vmin < threshold)) continue;
        double um[4];
        for (int v = 0; v < nvar; ++v) {
            for (int64_t q = ptr[e]; q < ptr[e + 1]; ++q)
                mean[v * N + e] = mean[v * N + e] + u[v * N + idx[q]];
            mean[v * N + e] = mean[v * N + e] / (double)(ptr[e + 1] - ptr[e]);
            um[v] = mean[v * N + e];
        }
        const double vmean = or_variable(variable, gamma, um);
        const double theta = (vmean - threshold) / (vmean - vmin);
        for (int v = 0; v < nvar; ++v)
            loc[v * N + e] = fma(theta, u[v * N + e], (1.0 - theta) * um[v]);
    }
    for (int64_t e = 0; e < N; ++e) {
        int nz = 0;
        for (int v = 0; v < nvar; ++v) nz |= (loc[v * N + e] != 0.0);
        if (nz)
            for (int v = 0; v < nvar; ++v) u[v * N + e] = loc[v * N + e];
    }
}

/* ======================================================================
 * SSPRK33 (OrdinaryDiffEq, outside this tree): Shu-Osher form
 *   u1 = u_n + dt f(u_n)                  (FSAL: f(u_n) from the last step)
 *   u2 = (3 u_n + u1 + dt f(u1)) / 4      at t + dt
 *   u_{n+1} = (u_n + 2 u2 + 2 dt f(u2))/3 at t + dt/2; f(u_{n+1}) at t + dt
 * with a stage limiter after each stage, success_iter += 1 per step, and
 * the HistoryCallback (history.jl:43-75) after each step and at init.
 * fsal must hold f(u_n) on entry when *success_iter_io > 0 (a restart);
 * on a fresh start (*success_iter_io == -1) it is computed here.
 * ==================================================================== */
int or_ssprk33(or_rhs_cfg* c, const or_step_cfg* s, double t0, double dt, int nsteps,
               double* u, double* fsal, double* time_hist, double* sol_hist, double* approx_du,
               int* success_iter_io)
{
    const int64_t N = c->N;
    const int64_t nv = (int64_t)c->nvar * N;
    double* uprev = (double*)malloc(sizeof(double) * nv);
    double* k = (double*)malloc(sizeof(double) * nv);
    double* scr = (double*)malloc(sizeof(double) * or_rhs_scratch_len(c));
    double* lscr = (double*)malloc(sizeof(double) * 2 * nv);
    int si = *success_iter_io;
    double t = t0;
    if (si < 0) {
        si = 0;
        c->success_iter = 0;
        if (s->approx_order > 0)
            or_history_update(nv, s->approx_order, 0, t, u, time_hist, sol_hist, approx_du);
        or_rhs(c, t, u, fsal, NULL, NULL, scr);
    }
#define OR_LIMIT(buf)                                                                       \
    do {                                                                                    \
        if (s->limiter_on)                                                                  \
            for (int l = 0; l < s->lim_count; ++l)                                          \
                or_limiter_zhang_shu(N, c->nvar, c->gamma, s->ptr, s->idx,                  \
                                     s->lim_threshold[l], s->lim_variable[l], buf, lscr);   \
    } while (0)
    for (int n = 0; n < nsteps; ++n) {
        c->success_iter = si;
        memcpy(uprev, u, sizeof(double) * nv);
        for (int64_t q = 0; q < nv; ++q) u[q] = fma(dt, fsal[q], uprev[q]);
        OR_LIMIT(u);
        or_rhs(c, t + dt, u, k, NULL, NULL, scr);
        for (int64_t q = 0; q < nv; ++q) u[q] = fma(3.0, uprev[q], fma(dt, k[q], u[q])) / 4.0;
        OR_LIMIT(u);
        or_rhs(c, t + dt / 2.0, u, k, NULL, NULL, scr);
        for (int64_t q = 0; q < nv; ++q)
            u[q] = fma(2.0, u[q], fma(2.0 * dt, k[q], uprev[q])) / 3.0;
        OR_LIMIT(u);
        or_rhs(c, t + dt, u, fsal, NULL, NULL, scr);
        t = t + dt;
        si += 1;
        if (s->approx_order > 0)
            or_history_update(nv, s->approx_order, si, t, u, time_hist, sol_hist, approx_du);
    }
#undef OR_LIMIT
    c->success_iter = si;
    *success_iter_io = si;
    free(uprev);
    free(k);
    free(scr);
    free(lscr);
    return 0;
}

/* ======================================================================
 * Adjoint.  No counterpart in the reference (which has none); the vector-
 * Jacobian product of the map u -> du of or_rhs for the flux divergence
 * with boundary conditions and no sources:
 *   du = M ( -sum_d D_d F_d(P u) ),  P = boundary projection on u,
 *   M = boundary mask on du.
 *   ubar = P^T sum_d -J_d(P u)^T (D_d^T (M lam))
 * scratch: 3 * nvar * N doubles.
 * ==================================================================== */
static void or_flux_jt(int eq, const double* adv_a, double gamma, int orientation,
                       const double* u, const double* g, double* out)
{
    if (eq == OR_EQ_ADVECTION) {
        out[0] = adv_a[orientation] * g[0];
        return;
    }
    const double rho = u[0], m1 = u[1], m2 = u[2], E = u[3];
    const double v1 = m1 / rho, v2 = m2 / rho;
    const double gm = gamma - 1.0;
    const double q = 0.5 * (v1 * v1 + v2 * v2);
    const double p = gm * (E - rho * q);
    const double H = (E + p) / rho;
    double J[4][4];
    if (orientation == 0) {
        const double r0[4] = {0.0, 1.0, 0.0, 0.0};
        const double r1[4] = {gm * q - v1 * v1, (3.0 - gamma) * v1, -gm * v2, gm};
        const double r2[4] = {-v1 * v2, v2, v1, 0.0};
        const double r3[4] = {v1 * (gm * q - H), H - gm * v1 * v1, -gm * v1 * v2, gamma * v1};
        memcpy(J[0], r0, sizeof r0);
        memcpy(J[1], r1, sizeof r1);
        memcpy(J[2], r2, sizeof r2);
        memcpy(J[3], r3, sizeof r3);
    } else {
        const double r0[4] = {0.0, 0.0, 1.0, 0.0};
        const double r1[4] = {-v1 * v2, v2, v1, 0.0};
        const double r2[4] = {gm * q - v2 * v2, -gm * v1, (3.0 - gamma) * v2, gm};
        const double r3[4] = {v2 * (gm * q - H), -gm * v1 * v2, H - gm * v2 * v2, gamma * v2};
        memcpy(J[0], r0, sizeof r0);
        memcpy(J[1], r1, sizeof r1);
        memcpy(J[2], r2, sizeof r2);
        memcpy(J[3], r3, sizeof r3);
    }
    for (int a = 0; a < 4; ++a) {
        double acc = 0.0;
        for (int b = 0; b < 4; ++b) acc += J[b][a] * g[b];
        out[a] = acc;
    }
}

int or_rhs_vjp(const or_rhs_cfg* c, double t, const double* u_in, const double* lam,
               double* ubar, double* scratch)
{
    const int64_t N = c->N;
    const int nv = c->nvar;
    const int64_t NV = (int64_t)nv * N;
    double* u = scratch;            /* P u */
    double* ml = scratch + NV;      /* M lam */
    double* g = scratch + 2 * NV;   /* D_d^T (M lam), one dimension at a time */
    double* junk = (double*)malloc(sizeof(double) * NV);
    memcpy(u, u_in, sizeof(double) * NV);
    memset(junk, 0, sizeof(double) * NV);
    or_bc_pass(c, t, u, junk);
    memcpy(ml, lam, sizeof(double) * NV);
    {
        /* the mask M: same zeroing the boundary pass applies to du */
        double* tmpu = (double*)malloc(sizeof(double) * NV);
        memcpy(tmpu, u, sizeof(double) * NV);
        or_bc_pass(c, t, tmpu, ml);
        free(tmpu);
    }
    memset(ubar, 0, sizeof(double) * NV);
    const double* D[2] = {c->dx, c->dy};
    for (int d = 0; d < 2; ++d) {
        memset(g, 0, sizeof(double) * NV);
        for (int v = 0; v < nv; ++v)
            or_csc_tmul(N, c->colptr, c->rowval, D[d], ml + v * N, g + v * N, 1.0);
        for (int64_t e = 0; e < N; ++e) {
            double ue[4], ge[4], oe[4];
            for (int v = 0; v < nv; ++v) {
                ue[v] = u[v * N + e];
                ge[v] = g[v * N + e];
            }
            or_flux_jt(c->eq, c->adv_a, c->gamma, d, ue, ge, oe);
            for (int v = 0; v < nv; ++v) ubar[v * N + e] -= oe[v];
        }
    }
    /* P^T: Dirichlet nodes get zero, slip nodes the (symmetric) projection */
    for (int gi = 0; gi < c->ntags; ++gi) {
        const int kind = c->tag_kind[gi];
        for (int32_t q = c->tag_ptr[gi]; q < c->tag_ptr[gi + 1]; ++q) {
            const int32_t b = c->bidx[q];
            if (kind == OR_BC_DIRICHLET) {
                for (int v = 0; v < nv; ++v) ubar[v * N + b] = 0.0;
            } else if (kind == OR_BC_SLIP) {
                const double n1 = c->bnormal[2 * q], n2 = c->bnormal[2 * q + 1];
                const double nrm = sqrt(n1 * n1 + n2 * n2);
                const double e1 = n1 / nrm, e2 = n2 / nrm;
                const double a1 = ubar[N + b], a2 = ubar[2 * N + b];
                const double s = a1 * e1 + a2 * e2;
                ubar[N + b] = a1 - s * e1;
                ubar[2 * N + b] = a2 - s * e2;
            }
        }
    }
    free(junk);
    return 0;
}
