// rf_operator.hip -- the matrix-free FP64 operator of the fiber system on
// gfx950: y = A x, the right-hand side b, the residual b - A x and the
// velocity sums, without storing A or W.
//
// mat_prod of fortran/matrix_assembly.f90:159-265 (numeric inner integrals
// G_f_GQ, fortran/quadrature.f90:166-223), extended to the analytic inner
// integrals (fortran/Int_analytical.f90:5-293,663-767) and to the force term
// of the right-hand side (fortran/matrix_assembly.f90:632-905) and of the
// velocities (fortran/coeffs_and_vels.f90:543-746).
//
// For target fiber i and Legendre index k = 0..M the kernel forms
//     IF_k(i) = sum_{j != i} sum_p w_p P_k(s_p) g_j(x_i + s_p t_i),
//     g_j(X)  = sum_{l=0..M} G_l(X; j) a_{j,l},
// the inner integral over fiber j of the kernel times j's force density
// f_j(s) = sum_l a_{j,l} P_l(s) (a_{j,0} = F_j/2 or 0, a_{j,l} = the
// coefficients x).  Both sums are linear, so thread (target i, point p) adds
// g_j(x_i + s_p t_i) over all sources j in three registers, and the
// contraction over p with w_p P_k happens once per target at the end.
// Sources are read with wavefront-uniform addresses (scalar loads).
//
// Analytic: G_l = al_l I + be_l R R^T + ga_l (t R^T + R t^T) + de_l t t^T with
// al_l, ..., de_l Legendre combinations of the moments I_m, J_m, S_m.  Moving
// the Legendre coefficients onto the source vector (c_m = sum_l C[l][m] a_l,
// the monomial coefficients of f_j) gives
//     g = sum_m (I_m + 2e^2 J_m) c_m + R sum_m [be_m (R.c_m) + ga_m (t.c_m)]
//                                    + t sum_m [ga_m (R.c_m) + de_m (t.c_m)],
// be_m = J_m - 6e^2 S_m, ga_m = 6e^2 S_{m+1} - J_{m+1}, de_m = J_{m+2} - 6e^2 S_{m+2}:
// about 30 flops per m on top of the moments.
// Numeric: g = sum_q [ (1/|R| + 2e^2/|R|^3) f_q + (1/|R|^3 - 6e^2/|R|^5) R (R.f_q) ],
// R = X - x_j - s_q t_j, f_q = w_q f_j(s_q).
#include "rf_inner.h"
#include "rf_internal.h"
#include "rigidfibers.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace rf {
namespace {

struct OpKArgs {
    int N, LQ, T, src_per_split, src_stride;
    double eps;
    const double* pos;
    const double* orient;
    const double* src;  // per source fiber: [M+1][4] (analytic) or [LQ][4] (numeric)
    const double* wp;   // [M+1][LQ] w_p P_k(s_p)
    const double* pv;   // [LQ]
    double* part;       // [splits][N][M+1][3]
};

// Source vectors of the operator.  a_{j,0} (force term): mode 0 none; modes
// 1, 2 the right-hand side's F/2 (only F_z with analytic integrals,
// fortran/matrix_assembly.f90:857-859); mode 3 F/2.  a_{j,l}, l >= 1: the
// coefficients x (none in mode 2).  The host turns the mode into the masks
// fmask (force components on) and x_on.
__global__ void op_source_kernel(int N, int M, int LQ, int analytic, int fmask, int x_on,
                                 const double* __restrict__ force, const double* __restrict__ x,
                                 const double* __restrict__ orient, const double* __restrict__ wp,
                                 double* __restrict__ src, int stride)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = analytic ? (M + 1) : LQ;
    if (idx >= N * per) return;
    const int j = idx / per, r = idx - j * per;
    double a[kMaxTerms + 1][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) a[0][c] = ((fmask >> c) & 1) ? 0.5 * force[3 * j + c] : 0.0;
    for (int l = 1; l <= M; ++l)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[l][c] = x_on ? x[(size_t)j * 3 * M + (l - 1) * 3 + c] : 0.0;
    double v[3] = {0.0, 0.0, 0.0};
    if (analytic) {
        // monomial coefficient m = r of f_j(s) = sum_l a_l P_l(s)
        for (int l = r; l <= M; ++l) {
            const double cf = kLeg[l][r];
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] += cf * a[l][c];
        }
    } else {
        // w_q f_j(s_q), q = r
        for (int l = 0; l <= M; ++l) {
            const double w = wp[l * LQ + r];
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] += w * a[l][c];
        }
    }
    double* o = src + (size_t)j * stride + 4 * r;
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
    o[3] = analytic ? (orient[3 * j] * v[0] + orient[3 * j + 1] * v[1] + orient[3 * j + 2] * v[2]) : 0.0;
}

template <int M, bool ANALYTIC>
__global__ void __launch_bounds__(512) operator_kernel(const OpKArgs a)
{
    constexpr int NL = M + 1;
    extern __shared__ double sg[];  // [T][LQ][3]
    const int T = a.T, LQ = a.LQ, N = a.N;
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * T;
    const int jb = blockIdx.y * a.src_per_split;
    const int je = min(N, jb + a.src_per_split);
    const double e2 = 2.0 * a.eps * a.eps, e6 = 6.0 * a.eps * a.eps;
    const int t1 = tid / LQ, p1 = tid - t1 * LQ;
    const int i1 = i0 + t1;
    const bool on = (t1 < T) && (i1 < N);
    double X0 = 0, X1 = 0, X2 = 0;
    if (on) {
        const double s = a.pv[p1];
        X0 = a.pos[3 * i1 + 0] + s * a.orient[3 * i1 + 0];
        X1 = a.pos[3 * i1 + 1] + s * a.orient[3 * i1 + 1];
        X2 = a.pos[3 * i1 + 2] + s * a.orient[3 * i1 + 2];
    }
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int j = jb; j < je; ++j) {
        if (!on || j == i1) continue;
        const double xj[3] = {a.pos[3 * j], a.pos[3 * j + 1], a.pos[3 * j + 2]};
        const double tj[3] = {a.orient[3 * j], a.orient[3 * j + 1], a.orient[3 * j + 2]};
        const double* cs = a.src + (size_t)j * a.src_stride;
        if constexpr (ANALYTIC) {
            Moments<M> mo;
            analytic_moments<M, double>(X0, X1, X2, xj, tj, mo);
            double ax = 0.0, ay = 0.0, az = 0.0, sR = 0.0, sT = 0.0;
#pragma unroll
            for (int m = 0; m <= M; ++m) {
                const double cx = cs[4 * m], cy = cs[4 * m + 1], cz = cs[4 * m + 2], tc = cs[4 * m + 3];
                const double al = mo.I[m] + e2 * mo.J[m];
                ax += al * cx;
                ay += al * cy;
                az += al * cz;
                const double rc = mo.r0x * cx + mo.r0y * cy + mo.r0z * cz;
                const double be = mo.J[m] - e6 * mo.S[m];
                const double ga = e6 * mo.S[m + 1] - mo.J[m + 1];
                const double de = mo.J[m + 2] - e6 * mo.S[m + 2];
                sR += be * rc + ga * tc;
                sT += ga * rc + de * tc;
            }
            g0 += ax + sR * mo.r0x + sT * tj[0];
            g1 += ay + sR * mo.r0y + sT * tj[1];
            g2 += az + sR * mo.r0z + sT * tj[2];
        } else {
            const double bx = X0 - xj[0], by = X1 - xj[1], bz = X2 - xj[2];
            double h0 = 0.0, h1 = 0.0, h2 = 0.0;
            for (int q = 0; q < LQ; ++q) {
                const double s = a.pv[q];
                const double fx = cs[4 * q], fy = cs[4 * q + 1], fz = cs[4 * q + 2];
                const double rx = bx - tj[0] * s, ry = by - tj[1] * s, rz = bz - tj[2] * s;
                const double r2 = rx * rx + ry * ry + rz * rz;
                const double inv = 1.0 / sqrt(r2);
                const double inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
                const double dg = inv + e2 * inv3, of = inv3 - e6 * inv5;
                const double rf = of * (rx * fx + ry * fy + rz * fz);
                h0 += dg * fx + rf * rx;
                h1 += dg * fy + rf * ry;
                h2 += dg * fz + rf * rz;
            }
            g0 += h0;
            g1 += h1;
            g2 += h2;
        }
    }
    if (t1 < T) {
        sg[(t1 * LQ + p1) * 3 + 0] = g0;
        sg[(t1 * LQ + p1) * 3 + 1] = g1;
        sg[(t1 * LQ + p1) * 3 + 2] = g2;
    }
    __syncthreads();
    // IF_k(i) partial of this source split: sum_p w_p P_k(s_p) g(p)
    for (int it = tid; it < T * NL * 3; it += blockDim.x) {
        const int t = it / (NL * 3), rem = it - t * (NL * 3);
        const int k = rem / 3, c = rem - k * 3;
        const int i = i0 + t;
        if (i >= N) continue;
        double acc = 0.0;
        for (int p = 0; p < LQ; ++p) acc += a.wp[k * LQ + p] * sg[(t * LQ + p) * 3 + c];
        a.part[((size_t)blockIdx.y * N + i) * (NL * 3) + rem] = acc;
    }
}

// Sums the split partials in a fixed order and applies the output map of the
// mode (one thread per (fiber, k)):
//   mode 0: y = x + P(IF)       (A x)
//   mode 1: y = -(x + P(IF))    (b - A x; the force term is in IF)
//   mode 2: y = -P(IF)          (b)
//   mode 3: y6[6i + 3k + c] = IF_k(i)_c for k = 0, 1  (the velocity sums)
// with P_k(IF) = D1 t (t.IF_1) for k = 1 and gamma_k (IF_k - E_k t (t.IF_k))
// for k >= 2 (fortran/matrix_assembly.f90:341-395).
struct OpEpiArgs {
    int N, M, splits, mode;
    int use_x;    // modes 0, 1: the identity term x
    double sign;  // +1 mode 0, -1 modes 1, 2
    double D1;
    double Ek[kMaxTerms + 1], gam[kMaxTerms + 1];
    const double* orient;
    const double* part;
    const double* x;
    double* y;
};

__global__ void operator_epilogue_kernel(const OpEpiArgs a)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int NL = a.M + 1;
    if (idx >= a.N * NL) return;
    const int i = idx / NL, k = idx - i * NL;
    const size_t stride = (size_t)a.N * NL * 3;
    double s[3] = {0.0, 0.0, 0.0};
    const double* pp = a.part + (size_t)i * NL * 3 + 3 * k;
    for (int sp = 0; sp < a.splits; ++sp) {
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += pp[sp * stride + c];
    }
    if (a.mode == 3) {
        if (k <= 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) a.y[6 * (size_t)i + 3 * k + c] = s[c];
        }
        return;
    }
    if (k == 0) return;
    const double t0 = a.orient[3 * i], t1 = a.orient[3 * i + 1], t2 = a.orient[3 * i + 2];
    const double td = t0 * s[0] + t1 * s[1] + t2 * s[2];
    const double tt[3] = {t0, t1, t2};
    const size_t row = (size_t)i * 3 * a.M + 3 * (k - 1);
    // One arithmetic form for modes 0..2 (y = sign (x + P) with x = 0 in
    // mode 2): a three-way branch on the mode here was compiled to a store
    // of an undefined register for mode 2.
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double pc = (k == 1) ? a.D1 * td * tt[c] : a.gam[k] * (s[c] - a.Ek[k] * td * tt[c]);
        const double xv = a.use_x ? a.x[row + c] : 0.0;
        a.y[row + c] = a.sign * (xv + pc);
    }
}

template <int M>
hipError_t launch_op_m(const OpKArgs& k, dim3 grid, int threads, size_t lds, bool analytic, hipStream_t s)
{
    if (analytic)
        hipLaunchKernelGGL((operator_kernel<M, true>), grid, dim3(threads), lds, s, k);
    else
        hipLaunchKernelGGL((operator_kernel<M, false>), grid, dim3(threads), lds, s, k);
    return hipGetLastError();
}

}  // namespace

OperatorPlan plan_operator(int N, int M, int LQ, int analytic, int num_cus)
{
    OperatorPlan p;
    p.N = N;
    p.M = M;
    p.LQ = LQ;
    p.analytic = analytic;
    // T targets per workgroup, one thread per (target, point): prefer whole
    // 64-wide wavefronts and at most 512 threads.
    int T = 0;
    for (int t = 16; t >= 1; --t)
        if (t * LQ <= 512 && (t * LQ) % 64 == 0) {
            T = t;
            break;
        }
    if (T == 0) T = std::max(1, std::min(16, 256 / LQ));
    p.T = T;
    p.threads = T * LQ;
    const int tiles = (N + T - 1) / T;
    // enough workgroups for every CU several times over, with at least 8
    // sources per split
    const int want = (4 * num_cus + tiles - 1) / tiles;
    p.splits = std::max(1, std::min(want, std::max(1, N / 8)));
    p.src_per_split = (N + p.splits - 1) / p.splits;
    p.splits = (N + p.src_per_split - 1) / p.src_per_split;
    p.src_stride = 4 * (analytic ? (M + 1) : LQ);
    return p;
}

size_t operator_scratch_doubles(const OperatorPlan& p)
{
    return (size_t)p.N * p.src_stride + (size_t)p.splits * p.N * (p.M + 1) * 3;
}

hipError_t launch_operator(const OperatorPlan& p, const Tables& t, int mode, const double* pos, const double* orient,
                           const double* force, const double* x, const double* wp, const double* pv, double* scratch,
                           double* y, hipStream_t s)
{
    if (mode < 0 || mode > 3 || p.M != t.M || p.LQ != t.LQ || p.threads > 512) return hipErrorInvalidValue;
    double* src = scratch;
    double* part = scratch + (size_t)p.N * p.src_stride;
    {
        const int per = p.analytic ? (p.M + 1) : p.LQ;
        const int tot = p.N * per;
        // force term: none for A x; F_z only for the analytic right-hand side
        const int fmask = (mode == 0) ? 0 : ((mode == 1 || mode == 2) && p.analytic) ? 4 : 7;
        const int x_on = (mode == 2) ? 0 : 1;
        hipLaunchKernelGGL(op_source_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, p.N, p.M, p.LQ, p.analytic,
                           fmask, x_on, force, x, orient, wp, src, p.src_stride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    OpKArgs k;
    k.N = p.N;
    k.LQ = p.LQ;
    k.T = p.T;
    k.src_per_split = p.src_per_split;
    k.src_stride = p.src_stride;
    k.eps = t.eps;
    k.pos = pos;
    k.orient = orient;
    k.src = src;
    k.wp = wp;
    k.pv = pv;
    k.part = part;
    const size_t lds = sizeof(double) * (size_t)p.T * p.LQ * 3;
    const dim3 grid((p.N + p.T - 1) / p.T, p.splits);
    hipError_t e = hipErrorInvalidValue;
    switch (p.M) {
    case 1: e = launch_op_m<1>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    case 2: e = launch_op_m<2>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    case 3: e = launch_op_m<3>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    case 4: e = launch_op_m<4>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    case 5: e = launch_op_m<5>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    case 6: e = launch_op_m<6>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    case 7: e = launch_op_m<7>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    case 8: e = launch_op_m<8>(k, grid, p.threads, lds, p.analytic != 0, s); break;
    }
    if (e != hipSuccess) return e;
    return launch_operator_epilogue(p, t, mode, orient, x, scratch, y, s);
}

hipError_t launch_operator_epilogue(const OperatorPlan& p, const Tables& t, int mode, const double* orient,
                                    const double* x, const double* scratch, double* y, hipStream_t s)
{
    if (mode < 0 || mode > 3 || p.M != t.M) return hipErrorInvalidValue;
    const double* part = scratch + (size_t)p.N * p.src_stride;
    OpEpiArgs ea;
    ea.N = p.N;
    ea.M = p.M;
    ea.splits = p.splits;
    ea.mode = mode;
    ea.use_x = (mode == 0 || mode == 1) ? 1 : 0;
    ea.sign = (mode == 0) ? 1.0 : -1.0;
    ea.D1 = t.D1;
    for (int i = 0; i <= kMaxTerms; ++i) {
        ea.Ek[i] = t.Ek[i];
        ea.gam[i] = t.gammak[i];
    }
    ea.orient = orient;
    ea.part = part;
    ea.x = x;
    ea.y = y;
    const int tot = p.N * (p.M + 1);
    hipLaunchKernelGGL(operator_epilogue_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, ea);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The operator on caller-provided device buffers without any allocation
// (rf_operator_noalloc).  Its scratch -- the quadrature tables, the
// velocity sums and the operator's source vectors and split partials -- is
// one file-scope device array per GPU, resolved once per device with
// hipGetSymbolAddress; host code uses only that address.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocScratchDoubles = size_t(1) << 22;  // 32 MiB per GPU
constexpr size_t kTableDoubles = 1024;                      // wp[(M+1) LQ] (<= 864), pv[LQ] at 864
constexpr size_t kPvOffset = (kMaxTerms + 1) * kMaxQuad;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {nullptr};

// up to 128 table values per launch, passed by value
struct TableChunk {
    int offset, count;
    double v[128];
};

__global__ void store_table_kernel(double* __restrict__ dst, const TableChunk c)
{
    const int i = threadIdx.x;
    if (i < c.count) dst[c.offset + i] = c.v[i];
}

hipError_t store_tables(double* dst, const Tables& t, hipStream_t s)
{
    const int nwp = (t.M + 1) * t.LQ;
    auto put = [&](const double* src, int count, int offset) -> hipError_t {
        for (int b = 0; b < count; b += 128) {
            TableChunk c;
            c.offset = offset + b;
            c.count = std::min(128, count - b);
            for (int i = 0; i < c.count; ++i) c.v[i] = src[b + i];
            hipLaunchKernelGGL(store_table_kernel, dim3(1), dim3(128), 0, s, dst, c);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    hipError_t e = put(t.wp, nwp, 0);
    if (e == hipSuccess) e = put(t.pv, t.LQ, (int)kPvOffset);
    return e;
}

}  // namespace

// The device address of this GPU's scratch array (set-up once per device).
static int noalloc_scratch(double** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("rf_operator_noalloc: no current HIP device");
        return RF_EHIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("rf_operator_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return RF_EHIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return RF_OK;
}

}  // namespace rf

extern "C" int rf_operator_noalloc(const rf_params* params, int32_t num_fibers, int32_t mode, const double* pos,
                                   const double* orient, const double* force, const double* x, double* y,
                                   void* stream)
{
    using namespace rf;
    if (!params || !pos || !orient || !force || !y || (mode != RF_OP_RHS && !x) || mode < RF_OP_AX ||
        mode > RF_OP_VELOCITY) {
        set_error("rf_operator_noalloc: null buffer or unknown mode");
        return RF_EINVAL;
    }
    const rf_params& P = *params;
    if (num_fibers < 2 || P.num_terms < 1 || P.num_terms > kMaxTerms || P.num_quad_intervals < 1 ||
        3 * P.num_quad_intervals > kMaxQuad || !(P.slenderness > 0.0)) {
        set_error("rf_operator_noalloc: invalid parameters (need N>=2, 1<=M<=8, 1<=NoQI<=32, eps>0)");
        return RF_EINVAL;
    }
    Tables tab;
    build_tables(P.num_terms, P.num_quad_intervals, P.slenderness, &tab);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        set_error("rf_operator_noalloc: cannot query the device");
        return RF_EHIP;
    }
    const int N = num_fibers;
    OperatorPlan plan = plan_operator(N, tab.M, tab.LQ, P.use_analytical_integration ? 1 : 0, cus);
    // fewer source splits where the partials would not fit
    const size_t fixed = kTableDoubles + 6 * (size_t)N + (size_t)N * plan.src_stride;
    const size_t per_split = (size_t)N * (tab.M + 1) * 3;
    if (fixed + per_split > kNoAllocScratchDoubles) {
        set_error("rf_operator_noalloc: " + std::to_string(N) + " fibers need " +
                  std::to_string((fixed + per_split) * sizeof(double)) + " bytes of scratch; the maximum is " +
                  std::to_string(kNoAllocScratchDoubles * sizeof(double)) + " (rf_operator has no limit)");
        return RF_EINVAL;
    }
    const size_t max_splits = (kNoAllocScratchDoubles - fixed) / per_split;
    if ((size_t)plan.splits > max_splits) {
        plan.splits = (int)max_splits;
        plan.src_per_split = (N + plan.splits - 1) / plan.splits;
        plan.splits = (N + plan.src_per_split - 1) / plan.src_per_split;
    }
    double* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    double* wp = base;
    double* pv = base + kPvOffset;
    double* y6 = base + kTableDoubles;
    double* work = y6 + 6 * (size_t)N;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = store_tables(base, tab, s);
    const int kmode = (mode == RF_OP_AX) ? kOpAx : (mode == RF_OP_RESIDUAL) ? kOpResidual
                                                : (mode == RF_OP_RHS) ? kOpRhs : kOpVelocity;
    if (e == hipSuccess)
        e = launch_operator(plan, tab, kmode, pos, orient, force, x, wp, pv, work, (mode == RF_OP_VELOCITY) ? y6 : y,
                            s);
    if (e == hipSuccess && mode == RF_OP_VELOCITY)
        e = launch_velocity_epilogue(N, tab.d, tab.e, P.shear_rate, pos, orient, force, y6, y, y + 3 * (size_t)N, s);
    if (e != hipSuccess) {
        set_error(std::string("rf_operator_noalloc: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    return RF_OK;
}
