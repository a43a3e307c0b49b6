/*
 * oracle_check.c -- runs every entry point of the FP64 oracle
 * (oracle/rf_oracle.c) on small systems under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/sanitize/Makefile), and checks the
 * entry points against each other: matrix-free products and row subsets
 * against the assembled matrix, LU against GMRES, the time-stepping driver
 * against its own pieces.  Any sanitizer report aborts with a non-zero status.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rf_oracle.h"

static int g_fail = 0;

#define CHECK(cond)                                                               \
    do {                                                                          \
        if (!(cond)) {                                                            \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++g_fail;                                                             \
        }                                                                         \
    } while (0)

static double *dalloc(size_t n)
{
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

/* xorshift64*, uniform in [-1, 1) */
static double urand(unsigned long long *s)
{
    *s ^= *s >> 12;
    *s ^= *s << 25;
    *s ^= *s >> 27;
    return (double)((*s * 2685821657736338717ULL) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

/* N fibers on a jittered cubic grid of spacing 2.5 (fiber length 2). */
static void make_layout(int N, unsigned long long seed, double *pos, double *ori)
{
    int side = 1;
    while (side * side * side < N) ++side;
    for (int i = 0; i < N; ++i) {
        const int g[3] = {i % side, (i / side) % side, i / (side * side)};
        double nrm = 0.0;
        for (int c = 0; c < 3; ++c) {
            pos[3 * i + c] = 2.5 * g[c] + 0.2 * urand(&seed);
            ori[3 * i + c] = urand(&seed);
            nrm += ori[3 * i + c] * ori[3 * i + c];
        }
        nrm = sqrt(nrm);
        for (int c = 0; c < 3; ++c) ori[3 * i + c] /= nrm;
    }
}

static double max_abs_diff(const double *a, const double *b, size_t n)
{
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i] - b[i]));
    return m;
}

static double max_abs(const double *a, size_t n)
{
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
    return m;
}

static void check_system(int analytic, int N, int M, int noqi)
{
    const int LQ = 3 * noqi, n = 3 * M * N;
    const double eps = 0.01;
    double *pv = dalloc(LQ), *wv = dalloc(LQ), *lvec = dalloc((size_t)LQ * M);
    double *pos = dalloc(3 * N), *ori = dalloc(3 * N), *force = dalloc(3 * N);
    double *A = dalloc((size_t)n * n), *b = dalloc(n), *x = dalloc(n), *y = dalloc(n), *y2 = dalloc(n);
    double *vel = dalloc(3 * N), *rot = dalloc(3 * N), *vel2 = dalloc(3 * N), *rot2 = dalloc(3 * N);
    int *targets = (int *)calloc((size_t)N, sizeof(int));
    unsigned long long seed = 0x5EEDULL + 977ULL * (unsigned)N + (unsigned)analytic;

    orc_quad_pts_wts(noqi, pv, wv);
    CHECK(orc_legendre(M, LQ, pv, lvec) == 0);
    double wsum = 0.0;
    for (int q = 0; q < LQ; ++q) wsum += wv[q];
    CHECK(fabs(wsum - 2.0) < 1e-13); /* Gauss weights on [-1, 1] */

    make_layout(N, seed, pos, ori);
    for (int i = 0; i < N; ++i) force[3 * i + 2] = -1.0;
    orc_assemble_matrix(analytic, eps, N, M, pos, ori, LQ, pv, wv, lvec, A);
    orc_assemble_rhs(analytic, eps, N, M, pos, ori, force, LQ, pv, wv, lvec, b);
    for (int i = 0; i < n; ++i) x[i] = urand(&seed);

    /* y = A x by the matrix and by the block-by-block product */
    for (int r = 0; r < n; ++r) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) s += A[r + (size_t)c * n] * x[c];
        y[r] = s;
    }
    for (int i = 0; i < N; ++i) targets[i] = N - 1 - i;
    orc_apply_rows(analytic, eps, N, M, pos, ori, LQ, pv, wv, lvec, x, N, targets, y2);
    for (int t = 0; t < N; ++t) /* rows come back in target order */
        for (int k = 0; k < 3 * M; ++k)
            CHECK(fabs(y2[3 * M * t + k] - y[3 * M * targets[t] + k]) <= 1e-12 * (1.0 + max_abs(y, n)));
    if (!analytic) {
        orc_mat_prod(eps, N, M, pos, ori, LQ, pv, wv, lvec, x, y2);
        CHECK(max_abs_diff(y, y2, n) <= 1e-12 * (1.0 + max_abs(y, n)));
    }
    orc_assemble_rhs_rows(analytic, eps, N, M, pos, ori, force, LQ, pv, wv, lvec, N, targets, y2);
    for (int t = 0; t < N; ++t)
        for (int k = 0; k < 3 * M; ++k) CHECK(fabs(y2[3 * M * t + k] - b[3 * M * targets[t] + k]) <= 1e-13);

    /* GMRES against LU on the same system */
    double *Af = dalloc((size_t)n * n), *xl = dalloc(n);
    memcpy(Af, A, sizeof(double) * (size_t)n * n);
    memcpy(xl, b, sizeof(double) * n);
    CHECK(orc_lu_solve(n, Af, xl) == 0);
    memset(x, 0, sizeof(double) * n);
    int iters = 0;
    double be = 1.0;
    CHECK(orc_gmres(n, A, b, x, 10, 500, 1e-12, &iters, &be) == 0);
    CHECK(be <= 1e-12 && (iters > 0 || max_abs(b, n) == 0.0)); /* one fiber: b = 0 */
    CHECK(max_abs_diff(x, xl, n) <= 1e-9 * (1.0 + max_abs(xl, n)));
    memset(x, 0, sizeof(double) * n);
    if (N > 1) CHECK(orc_gmres(n, A, b, x, 2, 2, 1e-15, &iters, &be) == -4);

    /* velocities of a subset against all fibers, with a shear flow */
    orc_compute_velocities(analytic, eps, N, M, pos, ori, force, LQ, pv, wv, lvec, xl, vel, rot, 0.3);
    orc_compute_velocities_rows(analytic, eps, N, M, pos, ori, force, LQ, pv, wv, lvec, xl, N, targets, vel2,
                                rot2, 0.3);
    for (int t = 0; t < N; ++t)
        for (int c = 0; c < 3; ++c) {
            CHECK(fabs(vel2[3 * t + c] - vel[3 * targets[t] + c]) <= 1e-13);
            CHECK(fabs(rot2[3 * t + c] - rot[3 * targets[t] + c]) <= 1e-13);
        }

    /* two steps of the driver, direct and iterative; step 1's solution is
     * the LU solution above (same layout, orientations already unit) */
    const int steps = 2;
    double *xo = dalloc((size_t)steps * n), *bo = dalloc((size_t)steps * n);
    double *vo = dalloc(3 * N * steps), *ro = dalloc(3 * N * steps), *po = dalloc(3 * N * steps),
           *to = dalloc(3 * N * steps);
    int its[2] = {0, 0};
    for (int direct = 0; direct < 2; ++direct) {
        CHECK(orc_simulate(analytic, direct, eps, 0.1, steps, N, M, noqi, 10, 200, 1e-12, force, pos, ori, bo, xo,
                           vo, ro, po, to, its, 0.0) == 0);
        CHECK(max_abs_diff(bo, b, n) <= 1e-13);
        CHECK(max_abs_diff(xo, xl, n) <= 1e-9 * (1.0 + max_abs(xl, n)));
        for (size_t i = 0; i < (size_t)3 * N * steps; ++i) CHECK(isfinite(po[i]) && isfinite(to[i]));
    }
    /* a driver call with no outputs requested */
    CHECK(orc_simulate(analytic, 0, eps, 0.1, 1, N, M, noqi, 10, 200, 1e-10, force, pos, ori, NULL, NULL, NULL,
                       NULL, NULL, NULL, NULL, 0.0) == 0);

    free(Af), free(xl), free(xo), free(bo), free(vo), free(ro), free(po), free(to);
    free(pv), free(wv), free(lvec), free(pos), free(ori), free(force), free(A), free(b), free(x), free(y);
    free(y2), free(vel), free(rot), free(vel2), free(rot2), free(targets);
}

static void check_inner_integrals(void)
{
    /* analytic and numeric inner integrals agree for a well-separated pair;
     * the Legendre combinations run to their highest order (8) */
    const int noqi = 16, LQ = 3 * noqi, lmax = 8;
    double pv[48], wv[48], lvec[48 * 8];
    orc_quad_pts_wts(noqi, pv, wv);
    CHECK(orc_legendre(8, LQ, pv, lvec) == 0);
    const double xb[3] = {0.1, -0.2, 0.3}, pb[3] = {0.6, 0.0, 0.8}, xbar[3] = {2.0, 1.5, -1.0};
    double L[7 * 9], I[30], J[30], S[30];
    orc_analytical_int(xb, pb, xbar, lmax, L, I, J, S);
    for (int i = 0; i < 7 * 9; ++i) CHECK(isfinite(L[i]));
    for (int i = 0; i < 30; ++i) CHECK(isfinite(I[i]) && isfinite(J[i]) && isfinite(S[i]));
    for (int l = 0; l < 8; ++l) {
        double ga[6], gn[6];
        orc_G_analytic(xb, pb, xbar, 0.01, l, ga);
        orc_G_numeric(xb, pb, xbar, 0.01, LQ, pv, wv, lvec + (size_t)LQ * (l > 0 ? l - 1 : 0), gn);
        /* l = 1..5; above that the forward moment recurrences of the
         * analytic form lose digits at this distance (c < 10) */
        for (int c = 0; c < 6; ++c) {
            CHECK(isfinite(ga[c]) && isfinite(gn[c]));
            if (l >= 1 && l <= 5) CHECK(fabs(ga[c] - gn[c]) <= 1e-9);
        }
    }
    /* a point on the fiber's own line behind it (d = 0, the special case of
     * the J and S moments); in front of it the log of I_0 is singular */
    const double on_line[3] = {0.1 - 3.0 * 0.6, -0.2, 0.3 - 3.0 * 0.8};
    orc_analytical_int(xb, pb, on_line, lmax, L, I, J, S);
    for (int i = 0; i < 7 * 9; ++i) CHECK(isfinite(L[i]));
    /* a point far away takes the backward recurrences (c >= 10) */
    const double far[3] = {9.0, -7.0, 4.0};
    orc_analytical_int(xb, pb, far, lmax, L, I, J, S);
    for (int i = 0; i < 7 * 9; ++i) CHECK(isfinite(L[i]));
}

static void check_lu_singular(void)
{
    double A[9] = {1, 2, 3, 2, 4, 6, 0, 0, 0}; /* rank 1, last column zero */
    double b[3] = {1, 2, 3};
    CHECK(orc_lu_solve(3, A, b) > 0);
}

int main(void)
{
    CHECK(orc_num_threads() >= 1);
    check_inner_integrals();
    check_lu_singular();
    for (int analytic = 0; analytic < 2; ++analytic) {
        check_system(analytic, 1, 5, 8);
        check_system(analytic, 7, 5, 8);
        check_system(analytic, 9, 3, 4);
    }
    printf("oracle_check: %d failed checks\n", g_fail);
    return g_fail ? 1 : 0;
}
