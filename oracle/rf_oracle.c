/*
 * rf_oracle.c -- FP64 CPU restatement of the Fortran rigid-fiber timestep
 * (ericwolter/rigidfibers_kth, fortran/).  TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP product and the CPU baseline timed beside it; the
 * product never links it.
 *
 * Every function cites the Fortran routine it follows.  Floating-point
 * expressions keep the Fortran's operand order (built with
 * -ffp-contract=off) so the restatement agrees with the reference to
 * rounding.  The only structural change: the analytic moments of a
 * (pair, point) are computed once and shared by all Legendre indices l,
 * where the Fortran recomputes them per l (fortran/matrix_assembly.f90:522-547
 * calls Analytical_int for every l).  The recurrence values do not depend on
 * the upper index, so the results are value-identical.
 */
#include "rf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* Quadrature, Legendre, constants                                         */
/* ---------------------------------------------------------------------- */

/* fortran/quadrature.f90:44-84: composite 3-point Gauss-Legendre. */
void orc_quad_pts_wts(int noqi, double *pv, double *wv)
{
    double p0[3], w0[3];
    p0[0] = -sqrt(15.0) / 5.0;
    p0[1] = 0.0;
    p0[2] = sqrt(15.0) / 5.0;
    w0[0] = 5.0 / 9.0;
    w0[1] = 8.0 / 9.0;
    w0[2] = 5.0 / 9.0;
    double a = -1.0;
    double iv = 2.0 / noqi;
    for (int i = 0; i < noqi; ++i) {
        for (int q = 0; q < 3; ++q) {
            pv[3 * i + q] = (2.0 * a + iv + p0[q] * iv) / 2.0;
            wv[3 * i + q] = w0[q] / noqi;
        }
        a = a + iv;
    }
}

/* fortran/matrix_assembly.f90:36-56: explicit P_1..P_8. */
static double legendre_p(int k, double x)
{
    switch (k) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 0.5 * (3.0 * x * x - 1.0);
    case 3: return 0.5 * (5.0 * (x * x * x) - 3.0 * x);
    case 4: return 0.125 * (35.0 * ((x * x) * (x * x)) - 30.0 * (x * x) + 3.0);
    case 5: return 0.125 * (63.0 * (((x * x) * (x * x)) * x) - 70.0 * (x * x * x) + 15.0 * x);
    case 6: return 0.0625 * (231.0 * pow(x, 6) - 315.0 * pow(x, 4) + 105.0 * (x * x) - 5.0);
    case 7: return 0.0625 * (429.0 * pow(x, 7) - 693.0 * pow(x, 5) + 315.0 * pow(x, 3) - 35.0 * x);
    case 8: return 6435.0 / 128.0 * pow(x, 8) - 3003.0 / 32.0 * pow(x, 6) + 3465.0 / 64.0 * pow(x, 4)
                   - 315.0 / 32.0 * (x * x) + 35.0 / 128.0;
    }
    return NAN;
}

int orc_legendre(int M, int LQ, const double *pv, double *lvec)
{
    if (M < 1 || M > 8) return -1; /* matrix_assembly.f90:53-55 */
    for (int k = 1; k <= M; ++k)
        for (int p = 0; p < LQ; ++p) lvec[(k - 1) * LQ + p] = legendre_p(k, pv[p]);
    return 0;
}

/* fortran/matrix_assembly.f90:294-305 and :370 */
typedef struct {
    double d, e, D1;
    double lambda[9], Ek[9], gammak[9]; /* index k = 1..M */
} consts_t;

static void make_consts(double eps, int M, consts_t *k)
{
    double c = log(eps * eps * exp(1.0));
    k->d = -c;
    k->e = 2.0;
    const double cc = 1.0;
    k->D1 = 0.75 / (k->d - 2.0 * cc);
    k->lambda[1] = 2.0;
    for (int i = 2; i <= M; ++i) k->lambda[i] = k->lambda[i - 1] + 2.0 / i;
    for (int i = 1; i <= M; ++i) {
        k->Ek[i] = (k->d - k->e - cc * k->lambda[i]) / 2.0 / (k->d - cc * k->lambda[i]);
        k->gammak[i] = 0.5 * (2.0 * i + 1.0) / (k->d + k->e - cc * k->lambda[i]);
    }
}

/* ---------------------------------------------------------------------- */
/* Analytic inner integral                                                 */
/* ---------------------------------------------------------------------- */

/* Legendre-weighted combination used at fortran/Int_analytical.f90:213-288:
 * out = sum_m P_l[m] X[m + shift], expression order as written there. */
static double lcomb(int l, const double *X, int s)
{
    /* X is 0-based moment array; Fortran X(ii) == X[ii-1], ii = l+1. */
    switch (l) {
    case 0: return X[0 + s];
    case 1: return X[1 + s];
    case 2: return 0.5 * (3.0 * X[2 + s] - X[0 + s]);
    case 3: return 0.5 * (5.0 * X[3 + s] - 3.0 * X[1 + s]);
    case 4: return 0.125 * (35.0 * X[4 + s] - 30.0 * X[2 + s] + 3.0 * X[0 + s]);
    case 5: return 0.125 * (63.0 * X[5 + s] - 70.0 * X[3 + s] + 15.0 * X[1 + s]);
    case 6: return 0.0625 * (231.0 * X[6 + s] - 315.0 * X[4 + s] + 105.0 * X[2 + s] - 5.0 * X[0 + s]);
    case 7: return 0.0625 * (429.0 * X[7 + s] - 693.0 * X[5 + s] + 315.0 * X[3 + s] - 35.0 * X[1 + s]);
    case 8: return 1.0 / 128.0 * (6435.0 * X[8 + s] - 12012.0 * X[6 + s] + 6930.0 * X[4 + s]
                                  - 1260.0 * X[2 + s] + 35.0 * X[0 + s]);
    }
    return NAN;
}

/* fortran/Int_analytical.f90:5-293 (Analytical_int).  Moment arrays are
 * 0-based: I[m] = integral of s^m/|R| (Fortran I(m+1)). */
void orc_analytical_int(const double xb[3], const double pb[3], const double xbar[3],
                        int lmax, double *L, double *I, double *J, double *S)
{
    double Iw[32], Jw[32], Sw[32];
    for (int i = 0; i < 32; ++i) Iw[i] = Jw[i] = Sw[i] = 0.0;
    const double clim = 10.0;
    double R0[3] = {xbar[0] - xb[0], xbar[1] - xb[1], xbar[2] - xb[2]};
    double b = -2.0 * (pb[0] * R0[0] + pb[1] * R0[1] + pb[2] * R0[2]);
    double c = R0[0] * R0[0] + R0[1] * R0[1] + R0[2] * R0[2];
    const double se = 1.0, sb = -1.0;
    double ue = sqrt(1.0 + b + c);
    double ub = sqrt(1.0 - b + c);
    double te = 1.0 + 0.5 * b;
    double tb = -1.0 + 0.5 * b;
    double d = c - 0.25 * b * b;
    /* Fortran index ii (1-based) <-> C index ii-1 */
    Iw[0] = log(fabs(2.0 * se + b + 2.0 * ue)) - log(fabs(2.0 * sb + b + 2.0 * ub));
    Iw[1] = ue - ub - 0.5 * b * Iw[0];
    int nmax = lmax + 3; /* Fortran DO ii=3,N+3 */
    if (c < clim) {
        double nn = 2.0;
        for (int ii = 3; ii <= nmax; ++ii) {
            nn = nn + 1.0;
            double sgn = ((ii - 2) % 2 == 0) ? 1.0 : -1.0; /* sb**(ii-2) */
            Iw[ii - 1] = -(ue / (1.0 - nn)) + sgn * ub / (1.0 - nn)
                         - (0.5 - (nn - 1.0)) * b / (1.0 - nn) * Iw[ii - 2]
                         + (nn - 2.0) * c / (1.0 - nn) * Iw[ii - 3];
        }
    } else {
        Iw[29] = 0.0;
        Iw[28] = 0.0;
        double nn = 27.0;
        for (int ii = 28; ii >= 3; --ii) {
            double sgn = (((int)nn + 1) % 2 == 0) ? 1.0 : -1.0; /* sb**(nn+1) */
            Iw[ii - 1] = (nn + 2.0) / ((nn + 1.0) * c)
                         * (ue / (nn + 2.0) - sgn * ub / (nn + 2.0)
                            + (1.0 - 2.0 * (nn + 2.0)) / (2.0 * (nn + 2.0)) * b * Iw[ii]
                            - Iw[ii + 1]);
            nn = nn - 1.0;
        }
    }
    if (d < 1e-14) {
        Jw[0] = -1.0 / (2.0 * pow(se + b / 2.0, 2.0)) + 1.0 / (2.0 * pow(sb + b / 2.0, 2.0));
    } else {
        Jw[0] = te / (d * ue) - tb / (d * ub);
    }
    Jw[1] = -1.0 / ue + 1.0 / ub - b / 2.0 * Jw[0];
    if (c < clim) {
        for (int ii = 3; ii <= nmax; ++ii)
            Jw[ii - 1] = Iw[ii - 3] - b * Jw[ii - 2] - c * Jw[ii - 3];
    } else {
        Jw[29] = 0.0;
        Jw[28] = 0.0;
        for (int ii = 28; ii >= 3; --ii)
            Jw[ii - 1] = 1.0 / c * (Iw[ii - 1] - b * Jw[ii] - Jw[ii + 1]);
    }
    if (d < 1e-14) {
        Sw[0] = -4.0 / pow(2.0 * se + b, 4.0) + 4.0 / pow(2.0 * sb + b, 4.0);
        Sw[1] = -8.0 / (3.0 * pow(2.0 * se + b, 3.0)) + 8.0 / (3.0 * pow(2.0 * sb + b, 3.0))
                - b / 2.0 * Sw[0];
    } else {
        Sw[0] = (2.0 * se + b) / (6.0 * d * pow(ue, 3.0)) - (2.0 * sb + b) / (6.0 * d * pow(ub, 3.0))
                + 2.0 / (3.0 * d) * Jw[0];
        Sw[1] = -(b * se + 2.0 * c) / (6.0 * d * pow(ue, 3.0)) + (b * sb + 2.0 * c) / (6.0 * d * pow(ub, 3.0))
                - b / (3.0 * d) * Jw[0];
    }
    if (c < clim) {
        for (int ii = 3; ii <= nmax; ++ii)
            Sw[ii - 1] = Jw[ii - 3] - b * Sw[ii - 2] - c * Sw[ii - 3];
    } else {
        Sw[29] = 0.0;
        Sw[28] = 0.0;
        for (int ii = 28; ii >= 3; --ii)
            Sw[ii - 1] = 1.0 / c * (Jw[ii - 1] - b * Sw[ii] - Sw[ii + 1]);
    }
    int nl = lmax + 1;
    if (L) {
        for (int l = 0; l <= lmax; ++l) {
            L[0 * nl + l] = lcomb(l, Iw, 0);
            L[1 * nl + l] = lcomb(l, Jw, 0);
            L[2 * nl + l] = lcomb(l, Jw, 1);
            L[3 * nl + l] = lcomb(l, Jw, 2);
            L[4 * nl + l] = lcomb(l, Sw, 0);
            L[5 * nl + l] = lcomb(l, Sw, 1);
            L[6 * nl + l] = lcomb(l, Sw, 2);
        }
    }
    if (I) memcpy(I, Iw, 30 * sizeof(double));
    if (J) memcpy(J, Jw, 30 * sizeof(double));
    if (S) memcpy(S, Sw, 30 * sizeof(double));
}

/* Tensor of fortran/Int_analytical.f90:735-746 from one set of L-combos. */
static void g_tensor(const double R0[3], const double pb[3], double eps,
                     double L11, double L12, double L13, double L14,
                     double L22, double L23, double L24, double out6[6])
{
    static const int ia[6] = {0, 1, 2, 0, 0, 1};
    static const int ib[6] = {0, 1, 2, 1, 2, 2};
    double e2 = 2.0 * (eps * eps);
    for (int q = 0; q < 6; ++q) {
        int a = ia[q], bb = ib[q];
        double Id = (a == bb) ? 1.0 : 0.0;
        double R00 = R0[a] * R0[bb];
        double pR = pb[a] * R0[bb] + R0[a] * pb[bb]; /* pb_R_0 + R_0_pb */
        double pp = pb[a] * pb[bb];
        out6[q] = L11 * Id + R00 * L12 - pR * L13 + pp * L14
                  + e2 * (Id * L12 - 3.0 * R00 * L22 + 3.0 * pR * L23 - 3.0 * pp * L24);
    }
}

/* G for all l = 0..lmax at one point: G[6*l + comp]. */
static void g_analytic_all(const double xb[3], const double pb[3], const double xbar[3],
                           double eps, int lmax, double *G)
{
    double L[7 * 9];
    orc_analytical_int(xb, pb, xbar, lmax, L, NULL, NULL, NULL);
    int nl = lmax + 1;
    double R0[3] = {xbar[0] - xb[0], xbar[1] - xb[1], xbar[2] - xb[2]};
    for (int l = 0; l <= lmax; ++l)
        g_tensor(R0, pb, eps, L[l], L[nl + l], L[2 * nl + l], L[3 * nl + l], L[4 * nl + l],
                 L[5 * nl + l], L[6 * nl + l], G + 6 * l);
}

void orc_G_analytic(const double xb[3], const double pb[3], const double xbar[3],
                    double eps, int l, double out6[6])
{
    double G[6 * 9];
    g_analytic_all(xb, pb, xbar, eps, l, G);
    memcpy(out6, G + 6 * l, 6 * sizeof(double));
}

/* ---------------------------------------------------------------------- */
/* Numeric inner integral (fortran/quadrature.f90:100-223)                 */
/* ---------------------------------------------------------------------- */

/* Kernel components K11,K22,K33,K12,K13,K23 at R (quadrature.f90:137-148). */
static void stokes_kernel(double rx, double ry, double rz, double eps, double K[6])
{
    double rm = sqrt(rx * rx + ry * ry + rz * rz);
    double r3 = rm * rm * rm;
    double r5 = rm * rm * rm * rm * rm;
    double e2 = 2.0 * (eps * eps);
    double e6 = 6.0 * (eps * eps);
    K[0] = 1.0 / rm + rx * rx / r3 + e2 * (1.0 / r3 - 3.0 * (rx * rx) / r5);
    K[1] = 1.0 / rm + ry * ry / r3 + e2 * (1.0 / r3 - 3.0 * (ry * ry) / r5);
    K[2] = 1.0 / rm + rz * rz / r3 + e2 * (1.0 / r3 - 3.0 * (rz * rz) / r5);
    K[3] = rx * ry / r3 - e6 * rx * ry / r5;
    K[4] = rx * rz / r3 - e6 * rx * rz / r5;
    K[5] = ry * rz / r3 - e6 * ry * rz / r5;
}

void orc_G_numeric(const double xb[3], const double pb[3], const double xbar[3],
                   double eps, int LQ, const double *pv, const double *wv,
                   const double *lvec_m, double out6[6])
{
    for (int q = 0; q < 6; ++q) out6[q] = 0.0;
    for (int i = 0; i < LQ; ++i) {
        double rx = xbar[0] - (xb[0] + pb[0] * pv[i]);
        double ry = xbar[1] - (xb[1] + pb[1] * pv[i]);
        double rz = xbar[2] - (xb[2] + pb[2] * pv[i]);
        double K[6];
        stokes_kernel(rx, ry, rz, eps, K);
        for (int q = 0; q < 6; ++q) out6[q] += wv[i] * K[q] * lvec_m[i];
    }
}

/* quadrature.f90:166-223 (G_f_GQ) */
static void g_f_numeric(const double xb[3], const double pb[3], const double xbar[3],
                        double eps, int LQ, const double *fx, const double *fy,
                        const double *fz, const double *pv, const double *wv, double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < LQ; ++i) {
        double rx = xbar[0] - (xb[0] + pb[0] * pv[i]);
        double ry = xbar[1] - (xb[1] + pb[1] * pv[i]);
        double rz = xbar[2] - (xb[2] + pb[2] * pv[i]);
        double K[6];
        stokes_kernel(rx, ry, rz, eps, K);
        out[0] += wv[i] * (K[0] * fx[i] + K[3] * fy[i] + K[4] * fz[i]);
        out[1] += wv[i] * (K[3] * fx[i] + K[1] * fy[i] + K[5] * fz[i]);
        out[2] += wv[i] * (K[4] * fx[i] + K[5] * fy[i] + K[2] * fz[i]);
    }
}

/* ---------------------------------------------------------------------- */
/* Assembly                                                                */
/* ---------------------------------------------------------------------- */

/* Block writes of fortran/matrix_assembly.f90:345-395 for one (filno, fno, l)
 * given Gmat[p][6] for that l. */
static void write_block(double *A, int n, int M, int LQ, const double *wv, const double *lvec,
                        const consts_t *K, int filno, int fno, int l,
                        const double *Gmat, const double ta[3])
{
    /* thvec(q) = sum(wv*Gmat(:,q)*LvecMat(:,kk)) for all kk at once; each
     * entry still sums over p in increasing order. */
    double thall[8][6];
    for (int kk = 0; kk < M; ++kk)
        for (int q = 0; q < 6; ++q) thall[kk][q] = 0.0;
    for (int p = 0; p < LQ; ++p) {
        for (int q = 0; q < 6; ++q) {
            const double g = wv[p] * Gmat[6 * p + q];
            for (int kk = 0; kk < M; ++kk) thall[kk][q] += g * lvec[kk * LQ + p];
        }
    }
    for (int kk = 1; kk <= M; ++kk) {
        const double *th = thall[kk - 1];
        double Q1 = th[0] * ta[0] + th[3] * ta[1] + th[4] * ta[2];
        double Q2 = th[3] * ta[0] + th[1] * ta[1] + th[5] * ta[2];
        double Q3 = th[4] * ta[0] + th[5] * ta[1] + th[2] * ta[2];
        int rowno = filno * 3 * M + 3 * (kk - 1);
        int p = fno * M * 3 + 3 * (l - 1);
        double Q[3] = {Q1, Q2, Q3};
        if (kk == 1) {
            for (int a = 0; a < 3; ++a)
                for (int bb = 0; bb < 3; ++bb)
                    A[(size_t)(rowno + a) + (size_t)(p + bb) * n] = K->D1 * ta[a] * Q[bb];
        } else {
            static const int sym[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
            double Ek = K->Ek[kk], g = K->gammak[kk];
            for (int a = 0; a < 3; ++a)
                for (int bb = 0; bb < 3; ++bb)
                    A[(size_t)(rowno + a) + (size_t)(p + bb) * n] =
                        g * (th[sym[a][bb]] - Ek * ta[a] * Q[bb]);
        }
    }
}

void orc_assemble_matrix(int analytic, double eps, int N, int M,
                         const double *pos, const double *orient,
                         int LQ, const double *pv, const double *wv,
                         const double *lvec, double *A)
{
    /* Each (filno, fno) block depends only on that pair, so the blocks are
     * computed source-fiber (column-block) major here: a thread owns whole
     * columns of the column-major matrix.  Values are those of the Fortran's
     * filno-major loop (matrix_assembly.f90:499-626). */
    consts_t K;
    make_consts(eps, M, &K);
    int n = 3 * M * N;
#pragma omp parallel
    {
        double *Gall = (double *)malloc(sizeof(double) * LQ * 6 * (M + 1));
        double *Gmat = (double *)malloc(sizeof(double) * LQ * 6);
#pragma omp for schedule(dynamic, 1)
        for (int fno = 0; fno < N; ++fno) {
            /* zero this fiber's columns, identity on its diagonal block */
            for (int c = 3 * M * fno; c < 3 * M * (fno + 1); ++c) {
                double *col = A + (size_t)c * n;
                memset(col, 0, sizeof(double) * n);
                col[c] = 1.0;
            }
            const double *xcb = pos + 3 * fno, *tb = orient + 3 * fno;
            for (int filno = 0; filno < N; ++filno) {
                if (fno == filno) continue;
                const double *xc = pos + 3 * filno, *ta = orient + 3 * filno;
                if (analytic) {
                    for (int p = 0; p < LQ; ++p) {
                        double xbar[3] = {xc[0] + pv[p] * ta[0], xc[1] + pv[p] * ta[1],
                                          xc[2] + pv[p] * ta[2]};
                        g_analytic_all(xcb, tb, xbar, eps, M, Gall + (size_t)p * 6 * (M + 1));
                    }
                }
                for (int l = 1; l <= M; ++l) {
                    for (int p = 0; p < LQ; ++p) {
                        if (analytic) {
                            memcpy(Gmat + 6 * p, Gall + (size_t)p * 6 * (M + 1) + 6 * l, 6 * sizeof(double));
                        } else {
                            double xbar[3] = {xc[0] + pv[p] * ta[0], xc[1] + pv[p] * ta[1],
                                              xc[2] + pv[p] * ta[2]};
                            orc_G_numeric(xcb, tb, xbar, eps, LQ, pv, wv, lvec + (l - 1) * LQ, Gmat + 6 * p);
                        }
                    }
                    write_block(A, n, M, LQ, wv, lvec, &K, filno, fno, l, Gmat, ta);
                }
            }
        }
        free(Gall);
        free(Gmat);
    }
}

/* Rows of A x for the target fibers: the identity plus, against every
 * source fiber, the blocks of fortran/matrix_assembly.f90:345-395 (the same
 * G tensors and write_block as orc_assemble_matrix), without forming A.  A
 * checker of backward errors on systems too large for the dense matrix. */
void orc_apply_rows(int analytic, double eps, int N, int M,
                    const double *pos, const double *orient,
                    int LQ, const double *pv, const double *wv, const double *lvec,
                    const double *x, int ntargets, const int *targets, double *y)
{
    consts_t K;
    make_consts(eps, M, &K);
    const int m = 3 * M;
#pragma omp parallel
    {
        double *Gall = (double *)malloc(sizeof(double) * LQ * 6 * (M + 1));
        double *Gmat = (double *)malloc(sizeof(double) * LQ * 6);
        double *blk = (double *)malloc(sizeof(double) * m * m);
#pragma omp for schedule(dynamic, 1)
        for (int ti = 0; ti < ntargets; ++ti) {
            const int filno = targets[ti];
            const double *xc = pos + 3 * filno, *ta = orient + 3 * filno;
            double *yr = y + (size_t)m * ti;
            for (int r = 0; r < m; ++r) yr[r] = x[(size_t)filno * m + r];
            for (int fno = 0; fno < N; ++fno) {
                if (fno == filno) continue;
                const double *xcb = pos + 3 * fno, *tb = orient + 3 * fno;
                if (analytic) {
                    for (int p = 0; p < LQ; ++p) {
                        double xbar[3] = {xc[0] + pv[p] * ta[0], xc[1] + pv[p] * ta[1],
                                          xc[2] + pv[p] * ta[2]};
                        g_analytic_all(xcb, tb, xbar, eps, M, Gall + (size_t)p * 6 * (M + 1));
                    }
                }
                for (int l = 1; l <= M; ++l) {
                    for (int p = 0; p < LQ; ++p) {
                        if (analytic) {
                            memcpy(Gmat + 6 * p, Gall + (size_t)p * 6 * (M + 1) + 6 * l, 6 * sizeof(double));
                        } else {
                            double xbar[3] = {xc[0] + pv[p] * ta[0], xc[1] + pv[p] * ta[1],
                                              xc[2] + pv[p] * ta[2]};
                            orc_G_numeric(xcb, tb, xbar, eps, LQ, pv, wv, lvec + (l - 1) * LQ, Gmat + 6 * p);
                        }
                    }
                    /* the (filno, fno) block as a local m x m matrix */
                    write_block(blk, m, M, LQ, wv, lvec, &K, 0, 0, l, Gmat, ta);
                }
                const double *xs = x + (size_t)fno * m;
                for (int r = 0; r < m; ++r) {
                    double acc = 0.0;
                    for (int c = 0; c < m; ++c) acc += blk[r + (size_t)c * m] * xs[c];
                    yr[r] += acc;
                }
            }
        }
        free(Gall);
        free(Gmat);
        free(blk);
    }
}

void orc_assemble_rhs(int analytic, double eps, int N, int M,
                      const double *pos, const double *orient, const double *force,
                      int LQ, const double *pv, const double *wv,
                      const double *lvec, double *b)
{
    orc_assemble_rhs_rows(analytic, eps, N, M, pos, orient, force, LQ, pv, wv, lvec, N, NULL, b);
}

void orc_assemble_rhs_rows(int analytic, double eps, int N, int M,
                           const double *pos, const double *orient, const double *force,
                           int LQ, const double *pv, const double *wv,
                           const double *lvec, int ntargets, const int *targets, double *b)
{
    consts_t K;
    make_consts(eps, M, &K);
    memset(b, 0, sizeof(double) * 3 * M * (size_t)ntargets);
#pragma omp parallel
    {
        double *Gmat = (double *)malloc(sizeof(double) * LQ * 3);
        double *fx = (double *)malloc(sizeof(double) * LQ * 3);
        double *fy = fx + LQ, *fz = fx + 2 * LQ;
#pragma omp for schedule(dynamic, 1)
        for (int ti = 0; ti < ntargets; ++ti) {
            const int filno = targets ? targets[ti] : ti;
            const double *xc = pos + 3 * filno, *ta = orient + 3 * filno;
            double *brow = b + 3 * M * (size_t)ti;
            for (int fno = 0; fno < N; ++fno) {
                if (fno == filno) continue;
                const double *xcb = pos + 3 * fno, *tb = orient + 3 * fno;
                const double *F = force + 3 * fno;
                for (int p = 0; p < LQ; ++p) {
                    double xbar[3] = {xc[0] + pv[p] * ta[0], xc[1] + pv[p] * ta[1], xc[2] + pv[p] * ta[2]};
                    if (analytic) {
                        /* matrix_assembly.f90:857-859: only the z component of
                         * the external force enters the analytic RHS. */
                        double ff[3] = {0.0 * F[0], 0.0 * F[1], 0.5 * F[2]};
                        double G[6 * 1];
                        double L[7];
                        orc_analytical_int(xcb, tb, xbar, 0, L, NULL, NULL, NULL);
                        double R0[3] = {xbar[0] - xcb[0], xbar[1] - xcb[1], xbar[2] - xcb[2]};
                        g_tensor(R0, tb, eps, L[0], L[1], L[2], L[3], L[4], L[5], L[6], G);
                        /* Int_analytical.f90:894-896 */
                        Gmat[3 * p + 0] = G[0] * ff[0] + G[3] * ff[1] + G[4] * ff[2];
                        Gmat[3 * p + 1] = G[3] * ff[0] + G[1] * ff[1] + G[5] * ff[2];
                        Gmat[3 * p + 2] = G[4] * ff[0] + G[5] * ff[1] + G[2] * ff[2];
                    } else {
                        for (int q = 0; q < LQ; ++q) {
                            fx[q] = 0.5 * F[0];
                            fy[q] = 0.5 * F[1];
                            fz[q] = 0.5 * F[2];
                        }
                        g_f_numeric(xcb, tb, xbar, eps, LQ, fx, fy, fz, pv, wv, Gmat + 3 * p);
                    }
                }
                for (int kk = 1; kk <= M; ++kk) {
                    double contr[3];
                    for (int i = 0; i < 3; ++i) {
                        double s = 0.0;
                        for (int p = 0; p < LQ; ++p) s += wv[p] * Gmat[3 * p + i] * lvec[(kk - 1) * LQ + p];
                        contr[i] = s;
                    }
                    double tdc = ta[0] * contr[0] + ta[1] * contr[1] + ta[2] * contr[2];
                    double *r = brow + 3 * (kk - 1);
                    if (kk == 1) {
                        for (int a = 0; a < 3; ++a) r[a] = r[a] - K.D1 * tdc * ta[a];
                    } else {
                        for (int a = 0; a < 3; ++a)
                            r[a] = r[a] - K.gammak[kk] * (contr[a] - K.Ek[kk] * tdc * ta[a]);
                    }
                }
            }
        }
        free(Gmat);
        free(fx);
    }
}

/* fortran/matrix_assembly.f90:123-157 (assemble_forces): f_j(s_q). */
static void assemble_forces(int N, int M, int LQ, const double *lvec, const double *force,
                            const double *coeff, double *Fx, double *Fy, double *Fz)
{
    for (int i = 0; i < N; ++i) {
        for (int q = 0; q < LQ; ++q) {
            Fx[i * LQ + q] = 0.5 * force[3 * i + 0];
            Fy[i * LQ + q] = 0.5 * force[3 * i + 1];
            Fz[i * LQ + q] = 0.5 * force[3 * i + 2];
        }
        for (int kk = 1; kk <= M; ++kk) {
            int ind = i * 3 * M + (kk - 1) * 3;
            for (int q = 0; q < LQ; ++q) {
                Fx[i * LQ + q] = Fx[i * LQ + q] + coeff[ind + 0] * lvec[(kk - 1) * LQ + q];
                Fy[i * LQ + q] = Fy[i * LQ + q] + coeff[ind + 1] * lvec[(kk - 1) * LQ + q];
                Fz[i * LQ + q] = Fz[i * LQ + q] + coeff[ind + 2] * lvec[(kk - 1) * LQ + q];
            }
        }
    }
}

void orc_mat_prod(double eps, int N, int M, const double *pos, const double *orient,
                  int LQ, const double *pv, const double *wv, const double *lvec,
                  const double *x, double *y)
{
    consts_t K;
    make_consts(eps, M, &K);
    int n = 3 * M * N;
    double *zero = (double *)calloc(3 * (size_t)N, sizeof(double));
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)N * LQ);
    double *Fx = F, *Fy = F + (size_t)N * LQ, *Fz = F + 2 * (size_t)N * LQ;
    assemble_forces(N, M, LQ, lvec, zero, x, Fx, Fy, Fz);
    memcpy(y, x, sizeof(double) * n);
#pragma omp parallel
    {
        double *Gmat = (double *)malloc(sizeof(double) * LQ * 3);
#pragma omp for schedule(dynamic, 1)
        for (int filno = 0; filno < N; ++filno) {
            const double *xc = pos + 3 * filno, *ta = orient + 3 * filno;
            for (int fno = 0; fno < N; ++fno) {
                if (fno == filno) continue;
                const double *xcb = pos + 3 * fno, *tb = orient + 3 * fno;
                for (int p = 0; p < LQ; ++p) {
                    double xbar[3] = {xc[0] + pv[p] * ta[0], xc[1] + pv[p] * ta[1], xc[2] + pv[p] * ta[2]};
                    g_f_numeric(xcb, tb, xbar, eps, LQ, Fx + fno * LQ, Fy + fno * LQ, Fz + fno * LQ,
                                pv, wv, Gmat + 3 * p);
                }
                for (int kk = 1; kk <= M; ++kk) {
                    double IF[3];
                    for (int i = 0; i < 3; ++i) {
                        double s = 0.0;
                        for (int p = 0; p < LQ; ++p) s += wv[p] * (Gmat[3 * p + i] * lvec[(kk - 1) * LQ + p]);
                        IF[i] = s;
                    }
                    double tdi = ta[0] * IF[0] + ta[1] * IF[1] + ta[2] * IF[2];
                    double *r = y + filno * 3 * M + 3 * (kk - 1);
                    for (int a = 0; a < 3; ++a) {
                        double aac = (kk == 1) ? K.D1 * tdi * ta[a]
                                               : K.gammak[kk] * (IF[a] - K.Ek[kk] * tdi * ta[a]);
                        r[a] = r[a] + aac;
                    }
                }
            }
        }
        free(Gmat);
    }
    free(zero);
    free(F);
}

/* ---------------------------------------------------------------------- */
/* Velocities (fortran/coeffs_and_vels.f90:543-746)                        */
/* ---------------------------------------------------------------------- */

void orc_compute_velocities(int analytic, double eps, int N, int M,
                            const double *pos, const double *orient, const double *force,
                            int LQ, const double *pv, const double *wv, const double *lvec,
                            const double *coeff, double *vel, double *rot, double sh)
{
    orc_compute_velocities_rows(analytic, eps, N, M, pos, orient, force, LQ, pv, wv, lvec, coeff,
                                N, NULL, vel, rot, sh);
}

void orc_compute_velocities_rows(int analytic, double eps, int N, int M,
                                 const double *pos, const double *orient, const double *force,
                                 int LQ, const double *pv, const double *wv, const double *lvec,
                                 const double *coeff, int ntargets, const int *targets,
                                 double *vel, double *rot, double sh)
{
    double c = log(eps * eps * exp(1.0));
    double d = -c, e = 2.0, mubar = d;
    double *Fall = (double *)malloc(sizeof(double) * 3 * (size_t)N * LQ);
    double *Fx = Fall, *Fy = Fall + (size_t)N * LQ, *Fz = Fall + 2 * (size_t)N * LQ;
    assemble_forces(N, M, LQ, lvec, force, coeff, Fx, Fy, Fz);
#pragma omp parallel
    {
        double *G0 = (double *)malloc(sizeof(double) * LQ * 3);
        double *Gall = (double *)malloc(sizeof(double) * 6 * (M + 1));
#pragma omp for schedule(dynamic, 1)
        for (int ti = 0; ti < ntargets; ++ti) {
            const int filno = targets ? targets[ti] : ti;
            const double *xc = pos + 3 * filno, *ta = orient + 3 * filno;
            const double *Fa = force + 3 * filno;
            double tdF = ta[0] * Fa[0] + ta[1] * Fa[1] + ta[2] * Fa[2];
            double V[3], R[3];
            for (int a = 0; a < 3; ++a) V[a] = 0.5 * ((d + e) * Fa[a] + (d - e) * tdF * ta[a]);
            /* background shear u = (sh y, 0, 0): fortran/coeffs_and_vels.f90:597-600 */
            V[0] = V[0] + sh * mubar * xc[1];
            for (int a = 0; a < 3; ++a) R[a] = -sh * mubar * ta[1] * ta[0] * ta[a];
            R[0] = R[0] + sh * mubar * ta[1];
            for (int fno = 0; fno < N; ++fno) {
                if (fno == filno) continue;
                const double *xcb = pos + 3 * fno, *tb = orient + 3 * fno;
                for (int p = 0; p < LQ; ++p) {
                    double xbar[3] = {xc[0] + pv[p] * ta[0], xc[1] + pv[p] * ta[1], xc[2] + pv[p] * ta[2]};
                    if (analytic) {
                        /* Int_analytical.f90:772-850 (G_compute_GQ_f_kg) */
                        g_analytic_all(xcb, tb, xbar, eps, M, Gall);
                        const double *F = force + 3 * fno;
                        double g[3];
                        g[0] = 0.5 * F[0] * Gall[0] + 0.5 * F[1] * Gall[3] + 0.5 * F[2] * Gall[4];
                        g[1] = 0.5 * F[0] * Gall[3] + 0.5 * F[1] * Gall[1] + 0.5 * F[2] * Gall[5];
                        g[2] = 0.5 * F[0] * Gall[4] + 0.5 * F[1] * Gall[5] + 0.5 * F[2] * Gall[2];
                        for (int l = 1; l <= M; ++l) {
                            const double *cf = coeff + fno * 3 * M + (l - 1) * 3;
                            const double *Gl = Gall + 6 * l;
                            g[0] = g[0] + cf[0] * Gl[0] + cf[1] * Gl[3] + cf[2] * Gl[4];
                            g[1] = g[1] + cf[0] * Gl[3] + cf[1] * Gl[1] + cf[2] * Gl[5];
                            g[2] = g[2] + cf[0] * Gl[4] + cf[1] * Gl[5] + cf[2] * Gl[2];
                        }
                        G0[3 * p + 0] = g[0];
                        G0[3 * p + 1] = g[1];
                        G0[3 * p + 2] = g[2];
                    } else {
                        g_f_numeric(xcb, tb, xbar, eps, LQ, Fx + fno * LQ, Fy + fno * LQ, Fz + fno * LQ,
                                    pv, wv, G0 + 3 * p);
                    }
                }
                double IF0[3], IF1[3];
                for (int i = 0; i < 3; ++i) {
                    double s0 = 0.0, s1 = 0.0;
                    for (int p = 0; p < LQ; ++p) {
                        s0 += wv[p] * G0[3 * p + i] * 1.0;
                        s1 += wv[p] * G0[3 * p + i] * lvec[p];
                    }
                    IF0[i] = s0;
                    IF1[i] = s1;
                }
                double tI = ta[0] * IF1[0] + ta[1] * IF1[1] + ta[2] * IF1[2];
                for (int a = 0; a < 3; ++a) {
                    V[a] = V[a] + 0.5 * IF0[a];
                    R[a] = R[a] + 1.5 * (IF1[a] - tI * ta[a]);
                }
            }
            for (int a = 0; a < 3; ++a) {
                vel[3 * ti + a] = V[a] / mubar;
                rot[3 * ti + a] = R[a] / mubar;
            }
        }
        free(G0);
        free(Gall);
    }
    free(Fall);
}

/* ---------------------------------------------------------------------- */
/* Solvers                                                                 */
/* ---------------------------------------------------------------------- */

int orc_lu_solve(int n, double *A, double *b)
{
    int info = 0;
    for (int k = 0; k < n; ++k) {
        /* pivot: first index of max |A(i,k)|, i >= k (idamax) */
        int piv = k;
        double amax = fabs(A[(size_t)k + (size_t)k * n]);
        for (int i = k + 1; i < n; ++i) {
            double v = fabs(A[(size_t)i + (size_t)k * n]);
            if (v > amax) {
                amax = v;
                piv = i;
            }
        }
        if (piv != k) {
            for (int j = 0; j < n; ++j) {
                double t = A[(size_t)k + (size_t)j * n];
                A[(size_t)k + (size_t)j * n] = A[(size_t)piv + (size_t)j * n];
                A[(size_t)piv + (size_t)j * n] = t;
            }
            double t = b[k];
            b[k] = b[piv];
            b[piv] = t;
        }
        double akk = A[(size_t)k + (size_t)k * n];
        if (akk == 0.0) {
            if (!info) info = k + 1;
            continue;
        }
        double r = 1.0 / akk;
        for (int i = k + 1; i < n; ++i) A[(size_t)i + (size_t)k * n] *= r;
#pragma omp parallel for schedule(static) if (n - k > 256)
        for (int j = k + 1; j < n; ++j) {
            double akj = A[(size_t)k + (size_t)j * n];
            if (akj == 0.0) continue;
            double *colj = A + (size_t)j * n;
            const double *colk = A + (size_t)k * n;
            for (int i = k + 1; i < n; ++i) colj[i] -= colk[i] * akj;
        }
    }
    if (info) return info;
    /* forward (unit L) then backward (U) */
    for (int j = 0; j < n; ++j) {
        double bj = b[j];
        const double *col = A + (size_t)j * n;
        for (int i = j + 1; i < n; ++i) b[i] -= col[i] * bj;
    }
    for (int j = n - 1; j >= 0; --j) {
        const double *col = A + (size_t)j * n;
        b[j] = b[j] / col[j];
        double bj = b[j];
        for (int i = 0; i < j; ++i) b[i] -= col[i] * bj;
    }
    return 0;
}

static void matvec(int n, const double *A, const double *x, double *y)
{
    /* y = A x (dgemv 'N'), column-major */
#pragma omp parallel for schedule(static) if (n > 2000)
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    int nt = orc_num_threads();
    if (n <= 2000 || nt == 1) {
        for (int j = 0; j < n; ++j) {
            double xj = x[j];
            const double *col = A + (size_t)j * n;
            for (int i = 0; i < n; ++i) y[i] += col[i] * xj;
        }
        return;
    }
    /* row-partitioned to keep threads independent */
#pragma omp parallel
    {
#ifdef _OPENMP
        int t = omp_get_thread_num(), T = omp_get_num_threads();
#else
        int t = 0, T = 1;
#endif
        int lo = (int)((long long)n * t / T), hi = (int)((long long)n * (t + 1) / T);
        for (int j = 0; j < n; ++j) {
            double xj = x[j];
            const double *col = A + (size_t)j * n;
            for (int i = lo; i < hi; ++i) y[i] += col[i] * xj;
        }
    }
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* blas_rot.f:117-148 (drotg) */
static void drotg(double da, double db, double *c, double *s)
{
    double roe = db;
    if (fabs(da) > fabs(db)) roe = da;
    double scale = fabs(da) + fabs(db);
    if (scale == 0.0) {
        *c = 1.0;
        *s = 0.0;
        return;
    }
    double r = scale * sqrt((da / scale) * (da / scale) + (db / scale) * (db / scale));
    r = copysign(1.0, roe) * r;
    *c = da / r;
    *s = db / r;
}

/* blas_rot.f:78-115 (drot, n=1) */
static void drot1(double *x, double *y, double c, double s)
{
    double t = c * (*x) + s * (*y);
    *y = c * (*y) - s * (*x);
    *x = t;
}

/* dGMRES.f:1133-1146: t = V(:,1:k) y (dgemv 'N'), xCurrent = x + t. */
static void vy_plus_x(int n, int k, const double *V, const double *y, const double *x, double *xc)
{
    for (int i = 0; i < n; ++i) xc[i] = 0.0;
    for (int j = 0; j < k; ++j) {
        const double *vj = V + (size_t)j * n;
        double yj = y[j];
        for (int i = 0; i < n; ++i) xc[i] += vj[i] * yj;
    }
    for (int i = 0; i < n; ++i) xc[i] = x[i] + xc[i];
}

int orc_gmres(int n, const double *A, const double *b, double *x,
              int m, int max_iters, double tol, int *iters, double *be_out)
{
    /* dGMRES.f:736-1495 with icntl(4)=0 (no preconditioner), icntl(5)=0 (MGS),
     * icntl(6)=1 (initial guess), icntl(8)=1 (true residual at restart),
     * cntl(2:5)=0 so sA=sPA=0 and sb=sPb=||b||. */
    double *V = (double *)malloc(sizeof(double) * (size_t)n * (m + 1));
    double *w = (double *)malloc(sizeof(double) * n);
    double *r0 = (double *)malloc(sizeof(double) * n);
    double *xc = (double *)malloc(sizeof(double) * n);
    double *H = (double *)calloc((size_t)(m + 1) * (m + 2), sizeof(double)); /* H(i,j) = H[i + j*(m+1)] */
    double *y = (double *)malloc(sizeof(double) * (m + 1));
    double *cs = (double *)malloc(sizeof(double) * (m + 1));
    double *sn = (double *)malloc(sizeof(double) * (m + 1));
#define HH(i, j) H[(i) + (size_t)(j) * (m + 1)]
    int status = 0, iterOut = 0;
    double bn = sqrt(dot(n, b, b));
    if (bn == 0.0) {
        for (int i = 0; i < n; ++i) x[i] = 0.0;
        *iters = 0;
        *be_out = 0.0;
        goto done;
    }
    double sb = bn; /* sPb */
    matvec(n, A, x, r0);
    for (int i = 0; i < n; ++i) r0[i] = b[i] - r0[i];
    double beta = sqrt(dot(n, r0, r0));
    if (beta == 0.0) {
        *iters = 0;
        *be_out = 0.0;
        goto done;
    }
    for (int i = 0; i < n; ++i) V[i] = r0[i] * (1.0 / beta);
    for (;;) {
        HH(0, m) = beta; /* H(1,m+1)=beta, rhs of the least squares */
        for (int j = 1; j <= m; ++j) HH(j, m) = 0.0;
        int jH;
        double bea = 0.0;
        for (jH = 0; jH < m; ++jH) {
            matvec(n, A, V + (size_t)jH * n, w);
            for (int j = 0; j <= jH; ++j) HH(j, jH) = 0.0;
            for (int j = 0; j <= jH; ++j) { /* MGS */
                double dVi = dot(n, V + (size_t)j * n, w);
                HH(j, jH) = HH(j, jH) + dVi;
                double aux = -1.0 * dVi;
                const double *vj = V + (size_t)j * n;
                for (int i = 0; i < n; ++i) w[i] += aux * vj[i];
            }
            double dnormw = sqrt(dot(n, w, w));
            HH(jH + 1, jH) = dnormw;
            if (jH + 1 < m) {
                double aux = 1.0 / dnormw;
                double *vn = V + (size_t)(jH + 1) * n;
                for (int i = 0; i < n; ++i) vn[i] = aux * w[i];
            }
            for (int j = 0; j < jH; ++j) drot1(&HH(j, jH), &HH(j + 1, jH), cs[j], sn[j]);
            drotg(HH(jH, jH), HH(jH + 1, jH), &cs[jH], &sn[jH]);
            drot1(&HH(jH, m), &HH(jH + 1, m), cs[jH], sn[jH]);
            drot1(&HH(jH, jH), &HH(jH + 1, jH), cs[jH], sn[jH]);
            HH(jH + 1, jH) = 0.0;
            double dnormres = fabs(HH(jH + 1, m));
            bea = dnormres / sb;
            int total = iterOut * m + (jH + 1);
            if (bea <= tol || total >= max_iters) {
                /* xCurrent = x + V y, y = H(1:jH,1:jH)^-1 g */
                int k = jH + 1;
                for (int i = 0; i < k; ++i) y[i] = HH(i, m);
                for (int i = k - 1; i >= 0; --i) {
                    y[i] = y[i] / HH(i, i);
                    for (int r = 0; r < i; ++r) y[r] -= HH(r, i) * y[i];
                }
                vy_plus_x(n, k, V, y, x, xc);
                matvec(n, A, xc, w);
                for (int i = 0; i < n; ++i) w[i] = b[i] - w[i];
                double trueNormRes = sqrt(dot(n, w, w));
                double be = trueNormRes / sb;
                if (be <= tol || total >= max_iters) {
                    memcpy(x, xc, sizeof(double) * n);
                    *iters = total;
                    *be_out = be;
                    status = (be <= tol) ? 0 : -4;
                    goto done;
                }
            }
        }
        iterOut++;
        /* restart: x = x + V y (if not already), true residual */
        {
            int k = m;
            if (bea > tol) {
                for (int i = 0; i < k; ++i) y[i] = HH(i, m);
                for (int i = k - 1; i >= 0; --i) {
                    y[i] = y[i] / HH(i, i);
                    for (int r = 0; r < i; ++r) y[r] -= HH(r, i) * y[i];
                }
                vy_plus_x(n, k, V, y, x, xc);
            }
            memcpy(x, xc, sizeof(double) * n);
            matvec(n, A, x, r0);
            for (int i = 0; i < n; ++i) r0[i] = b[i] - r0[i];
            beta = sqrt(dot(n, r0, r0));
            for (int i = 0; i < n; ++i) V[i] = r0[i] * (1.0 / beta);
        }
    }
done:
#undef HH
    free(V);
    free(w);
    free(r0);
    free(xc);
    free(H);
    free(y);
    free(cs);
    free(sn);
    return status;
}

/* ---------------------------------------------------------------------- */
/* Driver (fortran/advect_fibers.f90:332-340, 437-523)                     */
/* ---------------------------------------------------------------------- */

int orc_simulate(int analytic, int direct, double eps, double dt, int steps,
                 int N, int M, int noqi, int restart, int max_iters, double tol,
                 const double *force, const double *pos0, const double *orient0,
                 double *b_out, double *x_out, double *vel_out, double *rot_out,
                 double *pos_out, double *orient_out, int *iters_out, double sh)
{
    int LQ = 3 * noqi, n = 3 * M * N, N3 = 3 * N;
    double *pv = (double *)malloc(sizeof(double) * LQ);
    double *wv = (double *)malloc(sizeof(double) * LQ);
    double *lvec = (double *)malloc(sizeof(double) * LQ * M);
    orc_quad_pts_wts(noqi, pv, wv);
    if (orc_legendre(M, LQ, pv, lvec)) {
        free(pv); free(wv); free(lvec);
        return -1;
    }
    double *X[3], *T[3], *Vv[2], *Rv[2];
    for (int i = 0; i < 3; ++i) {
        X[i] = (double *)calloc(N3, sizeof(double));
        T[i] = (double *)calloc(N3, sizeof(double));
    }
    for (int i = 0; i < 2; ++i) {
        Vv[i] = (double *)calloc(N3, sizeof(double));
        Rv[i] = (double *)calloc(N3, sizeof(double));
    }
    double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *b = (double *)malloc(sizeof(double) * n);
    double *x = (double *)calloc(n, sizeof(double));
    double *guess = (double *)calloc(n, sizeof(double));
    int one = 0, two = 1, three = 2, oldv = 0, newv = 1;
    memcpy(X[three], pos0, sizeof(double) * N3);
    for (int i = 0; i < N; ++i) {
        const double *t = orient0 + 3 * i;
        double tm = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int a = 0; a < 3; ++a) T[three][3 * i + a] = t[a] / tm;
    }
    int rc = 0;
    for (int tt = 1; tt <= steps; ++tt) {
        orc_assemble_matrix(analytic, eps, N, M, X[three], T[three], LQ, pv, wv, lvec, A);
        orc_assemble_rhs(analytic, eps, N, M, X[three], T[three], force, LQ, pv, wv, lvec, b);
        if (b_out) memcpy(b_out + (size_t)(tt - 1) * n, b, sizeof(double) * n);
        int its = 0;
        if (direct) {
            memcpy(x, b, sizeof(double) * n);
            orc_lu_solve(n, A, x);
        } else {
            double be;
            memcpy(x, guess, sizeof(double) * n);
            int st = orc_gmres(n, A, b, x, restart, max_iters, tol, &its, &be);
            if (st == -4) { /* coeffs_and_vels.f90:331-335 fallback */
                memcpy(x, b, sizeof(double) * n);
                orc_lu_solve(n, A, x);
            }
        }
        if (iters_out) iters_out[tt - 1] = its;
        if (x_out) memcpy(x_out + (size_t)(tt - 1) * n, x, sizeof(double) * n);
        memcpy(guess, x, sizeof(double) * n);
        oldv = newv;
        newv = 1 - oldv;
        orc_compute_velocities(analytic, eps, N, M, X[three], T[three], force, LQ, pv, wv, lvec, x,
                               Vv[newv], Rv[newv], sh);
        if (vel_out) memcpy(vel_out + (size_t)(tt - 1) * N3, Vv[newv], sizeof(double) * N3);
        if (rot_out) memcpy(rot_out + (size_t)(tt - 1) * N3, Rv[newv], sizeof(double) * N3);
        one = two;
        two = three;
        three = 3 - (one + two); /* 0-based form of three=6-(one+two) */
        for (int i = 0; i < N3; ++i) {
            if (tt == 1) {
                X[three][i] = X[two][i] + dt * Vv[newv][i];
                T[three][i] = T[two][i] + dt * Rv[newv][i];
            } else {
                X[three][i] = 4.0 / 3.0 * X[two][i] - 1.0 / 3.0 * X[one][i]
                              + 2.0 / 3.0 * dt * (2.0 * Vv[newv][i] - Vv[oldv][i]);
                T[three][i] = 4.0 / 3.0 * T[two][i] - 1.0 / 3.0 * T[one][i]
                              + 2.0 / 3.0 * dt * (2.0 * Rv[newv][i] - Rv[oldv][i]);
            }
        }
        for (int i = 0; i < N; ++i) {
            double *t = T[three] + 3 * i;
            double tm = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            for (int a = 0; a < 3; ++a) t[a] = t[a] / tm;
        }
        if (pos_out) memcpy(pos_out + (size_t)(tt - 1) * N3, X[three], sizeof(double) * N3);
        if (orient_out) memcpy(orient_out + (size_t)(tt - 1) * N3, T[three], sizeof(double) * N3);
    }
    for (int i = 0; i < 3; ++i) {
        free(X[i]);
        free(T[i]);
    }
    for (int i = 0; i < 2; ++i) {
        free(Vv[i]);
        free(Rv[i]);
    }
    free(A); free(b); free(x); free(guess); free(pv); free(wv); free(lvec);
    return rc;
}
