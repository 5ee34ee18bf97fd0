/* A plain C99 client of libmfrbf: the C ABI (include/mfrbf.h) driven without
 * Python, the way a compiled host (the Julia engine's ccall, a C++ solver)
 * drives it, with every result checked against a closed form:
 *   kNN        each node first in its own stencil at distance 0, distances
 *              ascending, the k-th distance equal to a brute-force search;
 *   weights    r^3 + degree 2: Dx, Dy reproduce the derivatives of x^2, xy,
 *              y^2 (and of constants: zero row sums);
 *   rhs!       linear advection of a linear field, velocity (1, 1):
 *              du = -(u_x + u_y) at every node (no boundary conditions);
 *   SSPRK33    the same field after 20 graph-replayed steps: u0 - 3t, since
 *              every stage stays linear;
 *   Euler      a constant state: du = 0 (free-stream preservation).
 * Exit 0: all checks passed; 1: a check or call failed; 2: no device. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mfrbf.h"

#define CHECK(call)                                                                        \
    do {                                                                                   \
        int rc_ = (call);                                                                  \
        if (rc_) {                                                                         \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, mf_last_error(ctx));       \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

#define EXPECT(cond, ...)                     \
    do {                                      \
        if (!(cond)) {                        \
            fprintf(stderr, __VA_ARGS__);     \
            fputc('\n', stderr);              \
            return 1;                         \
        }                                     \
    } while (0)

static double jitter(unsigned* s)
{
    *s = *s * 1664525u + 1013904223u;
    return (double)(*s >> 8) / 16777216.0 - 0.5;
}

int main(void)
{
    enum { SIDE = 40, N = SIDE * SIDE, K = 15 };
    static double x[N], y[N], u[N], du[N], u0[N], dist[N * K], wx[N * K], wy[N * K];
    static double e[4][N], de[4][N];
    static int32_t idx[N * K];
    unsigned seed = 20261203u;
    for (int i = 0; i < N; ++i) {
        x[i] = ((i % SIDE) + 0.5 + 0.3 * jitter(&seed)) / SIDE;
        y[i] = ((i / SIDE) + 0.5 + 0.3 * jitter(&seed)) / SIDE;
    }
    mf_ctx* ctx = NULL;
    if (mf_ctx_create(0, &ctx)) {
        fprintf(stderr, "no device: %s\n", ctx ? mf_last_error(ctx) : "mf_ctx_create failed");
        return 2;
    }
    CHECK(mf_set_equations(ctx, MF_EQ_ADVECTION_2D, 1.0, 1.0));
    CHECK(mf_set_points(ctx, x, y, N));
    double dx_min = 0.0, dx_avg = 0.0;
    CHECK(mf_knn(ctx, K, idx, dist, &dx_min, &dx_avg));
    for (int i = 0; i < N; ++i) {
        EXPECT(idx[i * K] == i && dist[i * K] == 0.0, "node %d is not first in its stencil", i);
        for (int s = 1; s < K; ++s) EXPECT(dist[i * K + s] >= dist[i * K + s - 1], "stencil %d not ascending", i);
    }
    for (int i = 0; i < N; i += 97) {  /* the k-th neighbour by brute force (squared distances) */
        const int jk = idx[i * K + K - 1];
        const double dkx = x[jk] - x[i], dky = y[jk] - y[i], d2k = dkx * dkx + dky * dky;
        int below = 0;
        for (int j = 0; j < N; ++j) {
            const double dx = x[j] - x[i], dy = y[j] - y[i];
            below += dx * dx + dy * dy < d2k;
        }
        EXPECT(below <= K - 1, "node %d: %d points closer than the k-th neighbour", i, below);
        EXPECT(fabs(dist[i * K + K - 1] - sqrt(d2k)) <= 1e-15, "node %d: k-th distance %.17g, %.17g by hand", i,
               dist[i * K + K - 1], sqrt(d2k));
    }
    EXPECT(dx_min > 0.0 && dx_avg >= dx_min, "dx_min %g, dx_avg %g", dx_min, dx_avg);

    CHECK(mf_compute_operator(ctx, 3, 2, 1, MF_OP_FLUX, wx, wy));
    double werr = 0.0;
    for (int i = 0; i < N; ++i) {
        double s0x = 0, s0y = 0, sxx = 0, sxy_x = 0, sxy_y = 0, syy = 0, scale = 0;
        for (int s = 0; s < K; ++s) {
            const int j = idx[i * K + s];
            const double a = wx[i * K + s], b = wy[i * K + s];
            s0x += a;
            s0y += b;
            sxx += a * x[j] * x[j];
            sxy_x += a * x[j] * y[j];
            sxy_y += b * x[j] * y[j];
            syy += b * y[j] * y[j];
            scale = fmax(scale, fmax(fabs(a), fabs(b)));
        }
        const double err = fmax(fmax(fabs(s0x), fabs(s0y)),
                                fmax(fmax(fabs(sxx - 2 * x[i]), fabs(syy - 2 * y[i])),
                                     fmax(fabs(sxy_x - y[i]), fabs(sxy_y - x[i]))));
        werr = fmax(werr, err / scale);
    }
    EXPECT(werr < 1e-9, "weights miss the quadratic derivatives by %.2e of max|w|", werr);

    double* uf[1] = {u};
    double* duf[1] = {du};
    for (int i = 0; i < N; ++i) u[i] = u0[i] = 1.0 + x[i] + 2.0 * y[i];
    CHECK(mf_rhs_host(ctx, 0.0, MF_PREC_F64_REF, uf, duf));
    double rerr = 0.0;
    for (int i = 0; i < N; ++i) rerr = fmax(rerr, fabs(du[i] + 3.0));
    EXPECT(rerr < 1e-9, "advection of a linear field: du misses -3 by %.2e", rerr);

    const double dt = 0.1 * dx_min;
    const int steps = 20;
    const double* uc[1] = {u0};
    CHECK(mf_state_upload(ctx, MF_BUF_U, uc));
    CHECK(mf_ssprk33(ctx, 0.0, dt, steps, MF_PREC_F64, MF_STEP_GRAPH, 0));
    CHECK(mf_state_download(ctx, MF_BUF_U, uf));
    double serr = 0.0;
    for (int i = 0; i < N; ++i) serr = fmax(serr, fabs(u[i] - (u0[i] - 3.0 * steps * dt)));
    EXPECT(serr < 1e-10, "SSPRK33 of a linear field: off by %.2e", serr);

    CHECK(mf_set_equations(ctx, MF_EQ_EULER_2D, 1.4, 0.0));
    const double state[4] = {1.2, 0.3, -0.2, 2.5};
    double* ef[4] = {e[0], e[1], e[2], e[3]};
    double* def[4] = {de[0], de[1], de[2], de[3]};
    for (int v = 0; v < 4; ++v)
        for (int i = 0; i < N; ++i) e[v][i] = state[v];
    CHECK(mf_rhs_host(ctx, 0.0, MF_PREC_F64_REF, ef, def));
    double ferr = 0.0;
    for (int v = 0; v < 4; ++v)
        for (int i = 0; i < N; ++i) ferr = fmax(ferr, fabs(de[v][i]));
    EXPECT(ferr < 1e-10, "Euler free stream: |du| = %.2e", ferr);

    CHECK(mf_ctx_destroy(ctx));
    printf("mfrbf C client ok: %d nodes, weights %.1e, rhs %.1e, steps %.1e, free stream %.1e\n", N, werr, rerr,
           serr, ferr);
    return 0;
}
