// rf_inner.h -- device math shared by the assembly kernel and the
// matrix-free operator: the Legendre monomial coefficients, the analytic
// moments I, J, S of the inner integral (fortran/Int_analytical.f90:5-293),
// the G_l tensor built from them (fortran/Int_analytical.f90:663-767) and the
// numeric inner integrals (fortran/quadrature.f90:100-164).
#pragma once

#include "rf_internal.h"

#include <type_traits>

namespace rf {
namespace {

// Coefficients of P_l(s) = sum_m c[l][m] s^m (the combinations of
// fortran/Int_analytical.f90:213-288).
__device__ constexpr double kLeg[9][9] = {
    {1, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 1, 0, 0, 0, 0, 0, 0, 0},
    {-0.5, 0, 1.5, 0, 0, 0, 0, 0, 0},
    {0, -1.5, 0, 2.5, 0, 0, 0, 0, 0},
    {0.375, 0, -3.75, 0, 4.375, 0, 0, 0, 0},
    {0, 1.875, 0, -8.75, 0, 7.875, 0, 0, 0},
    {-0.3125, 0, 6.5625, 0, -19.6875, 0, 14.4375, 0, 0},
    {0, -2.1875, 0, 19.6875, 0, -43.3125, 0, 26.8125, 0},
    {0.2734375, 0, -9.84375, 0, 54.140625, 0, -93.84375, 0, 50.2734375},
};

// Symmetric 3x3 <-> 6-vector (11,22,33,12,13,23).
__device__ __forceinline__ int sym6(int a, int b)
{
    return (a == b) ? a : (a + b + 2);  // (0,1)->3 (0,2)->4 (1,2)->5
}

// Moments of the analytic inner integral at point X against fiber (xj,tj):
// I_n, J_n, S_n = integral over s in [-1,1] of s^n / |R|^{1,3,5}, n = 0..M+2,
// R = X - xj - s tj (fortran/Int_analytical.f90:5-207).  The closed forms of
// n = 0, 1 (logarithms, square roots, the 1/d terms) and the forward
// recurrences of the near field (c < 10) are FP64; RB is the arithmetic type
// of the 26-step backward recurrences of the far field (double, or float in
// mixed precision, whose error is then ~FP32 rounding of the moments).
template <int M>
struct Moments {
    double r0x, r0y, r0z, px, py, pz;
    double I[M + 3], J[M + 3], S[M + 3];
};

template <int M, typename RB>
__device__ __forceinline__ void analytic_moments(double X0, double X1, double X2, const double* xj, const double* tj,
                                                 Moments<M>& mo)
{
    constexpr int NM = M + 3;  // moments 0..M+2
    const double r0x = X0 - xj[0], r0y = X1 - xj[1], r0z = X2 - xj[2];
    const double px = tj[0], py = tj[1], pz = tj[2];
    const double b = -2.0 * (px * r0x + py * r0y + pz * r0z);
    const double c = r0x * r0x + r0y * r0y + r0z * r0z;
    const double ue = sqrt(1.0 + b + c);
    const double ub = sqrt(1.0 - b + c);
    const double d = c - 0.25 * b * b;
    double* I = mo.I;
    double* J = mo.J;
    double* S = mo.S;
    I[0] = log(fabs(2.0 + b + 2.0 * ue)) - log(fabs(-2.0 + b + 2.0 * ub));
    I[1] = ue - ub - 0.5 * b * I[0];
    const double iue = 1.0 / ue, iub = 1.0 / ub;
    if (d < 1e-14) {  // collinear (Int_analytical.f90:162-169,189-196)
        const double sp = 1.0 + 0.5 * b, sm = -1.0 + 0.5 * b;
        J[0] = -0.5 / (sp * sp) + 0.5 / (sm * sm);
        J[1] = -iue + iub - 0.5 * b * J[0];
        const double ap = 2.0 + b, am = -2.0 + b;
        S[0] = -4.0 / (ap * ap * ap * ap) + 4.0 / (am * am * am * am);
        S[1] = -8.0 / (3.0 * ap * ap * ap) + 8.0 / (3.0 * am * am * am) - 0.5 * b * S[0];
    } else {
        const double id = 1.0 / d;
        J[0] = ((1.0 + 0.5 * b) * iue - (-1.0 + 0.5 * b) * iub) * id;
        J[1] = -iue + iub - 0.5 * b * J[0];
        const double iue3 = iue * iue * iue, iub3 = iub * iub * iub;
        const double i6d = id * (1.0 / 6.0);
        S[0] = ((2.0 + b) * iue3 - (-2.0 + b) * iub3) * i6d + ((2.0 / 3.0) * id) * J[0];
        S[1] = (-(b + 2.0 * c) * iue3 + (2.0 * c - b) * iub3) * i6d - (b * id * (1.0 / 3.0)) * J[0];
    }
    if (c < 10.0) {
        // forward recurrences (Int_analytical.f90:123-136,172-176,197-200)
#pragma unroll
        for (int n = 2; n < NM; ++n) {
            const double sgn = (n % 2 == 1) ? 1.0 : -1.0;  // (-1)^(n-1)
            I[n] = (ue - sgn * ub) * (1.0 / n) - ((n - 0.5) / n) * b * I[n - 1] - ((n - 1.0) / n) * c * I[n - 2];
            J[n] = I[n - 2] - b * J[n - 1] - c * J[n - 2];
            S[n] = J[n - 2] - b * S[n - 1] - c * S[n - 2];
        }
    } else {
        // backward recurrences from n = 27 with zero start
        // (Int_analytical.f90:137-158,177-183,201-206)
        const RB ic = (RB)(1.0 / c), rb = (RB)b;
        const RB uem = (RB)(ue - ub), uep = (RB)(ue + ub);
        RB i1 = 0, i2 = 0, j1 = 0, j2 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int m = 27; m >= 2; --m) {
            const RB base = ((m + 1) % 2 == 0) ? uem : uep;  // ue - (-1)^(m+1) ub
            const RB im = ic * RB(1.0 / (m + 1.0)) * (base - RB(m + 1.5) * rb * i1 - RB(m + 2.0) * i2);
            const RB jm = ic * (im - rb * j1 - j2);
            const RB sm = ic * (jm - rb * s1 - s2);
            if (m < NM) {
                I[m] = (double)im;
                J[m] = (double)jm;
                S[m] = (double)sm;
            }
            i2 = i1;
            i1 = im;
            j2 = j1;
            j1 = jm;
            s2 = s1;
            s1 = sm;
        }
    }
    mo.r0x = r0x;
    mo.r0y = r0y;
    mo.r0z = r0z;
    mo.px = px;
    mo.py = py;
    mo.pz = pz;
}

// G_l from the moments: the Legendre combinations L11..L24 of
// Int_analytical.f90:213-288 and the tensor of :735-736, regrouped as
// G = al I + be R0 R0^T + ga (t R0^T + R0 t^T) + de t t^T.
template <int M>
__device__ __forceinline__ void analytic_Gl(const Moments<M>& mo, int l, double e2, double e6, double (&g)[6])
{
    double L11 = 0, L12 = 0, L13 = 0, L14 = 0, L22 = 0, L23 = 0, L24 = 0;
#pragma unroll
    for (int m = 0; m <= kMaxTerms; ++m) {
        if (m <= l && kLeg[l][m] != 0.0) {
            const double cf = kLeg[l][m];
            L11 += cf * mo.I[m];
            L12 += cf * mo.J[m];
            L13 += cf * mo.J[m + 1];
            L14 += cf * mo.J[m + 2];
            L22 += cf * mo.S[m];
            L23 += cf * mo.S[m + 1];
            L24 += cf * mo.S[m + 2];
        }
    }
    const double al = L11 + e2 * L12;
    const double be = L12 - e6 * L22;
    const double ga = e6 * L23 - L13;
    const double de = L14 - e6 * L24;
    const double rx = mo.r0x, ry = mo.r0y, rz = mo.r0z, px = mo.px, py = mo.py, pz = mo.pz;
    g[0] = al + be * rx * rx + 2.0 * ga * px * rx + de * px * px;
    g[1] = al + be * ry * ry + 2.0 * ga * py * ry + de * py * py;
    g[2] = al + be * rz * rz + 2.0 * ga * pz * rz + de * pz * pz;
    g[3] = be * rx * ry + ga * (px * ry + rx * py) + de * px * py;
    g[4] = be * rx * rz + ga * (px * rz + rx * pz) + de * px * pz;
    g[5] = be * ry * rz + ga * (py * rz + ry * pz) + de * py * pz;
}

// Numeric inner integrals (fortran/quadrature.f90:121-155): LQ-point sum of
// the kernel times w_q P_l(s_q) over fiber j.
template <int M, typename R>
__device__ __forceinline__ void numeric_G(double X0, double X1, double X2, const double* xj, const double* tj,
                                          double eps, int LQ, const double* __restrict__ pv,
                                          const double* __restrict__ wp, double (&G)[(M + 1) * 6])
{
    R acc[(M + 1) * 6];
#pragma unroll
    for (int q = 0; q < (M + 1) * 6; ++q) acc[q] = 0;
    const R e2 = R(2.0 * eps * eps), e6 = R(6.0 * eps * eps);
    const double bx = X0 - xj[0], by = X1 - xj[1], bz = X2 - xj[2];
    for (int q = 0; q < LQ; ++q) {
        const double s = pv[q];
        const R rx = (R)(bx - tj[0] * s), ry = (R)(by - tj[1] * s), rz = (R)(bz - tj[2] * s);
        const R r2 = rx * rx + ry * ry + rz * rz;
        const R inv = R(1) / sqrt(r2);
        const R inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
        const R dg = inv + e2 * inv3, of = inv3 - e6 * inv5;
        const R K0 = dg + of * rx * rx, K1 = dg + of * ry * ry, K2 = dg + of * rz * rz;
        const R K3 = of * rx * ry, K4 = of * rx * rz, K5 = of * ry * rz;
#pragma unroll
        for (int l = 0; l <= M; ++l) {
            const R w = (R)wp[l * LQ + q];
            acc[l * 6 + 0] += w * K0;
            acc[l * 6 + 1] += w * K1;
            acc[l * 6 + 2] += w * K2;
            acc[l * 6 + 3] += w * K3;
            acc[l * 6 + 4] += w * K4;
            acc[l * 6 + 5] += w * K5;
        }
    }
#pragma unroll
    for (int q = 0; q < (M + 1) * 6; ++q) G[q] = (double)acc[q];
}

}  // namespace
}  // namespace rf
