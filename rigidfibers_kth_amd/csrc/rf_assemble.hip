// rf_assemble.hip -- pairwise assembly of the rigid-fiber linear system on
// gfx950.
//
// For every ordered pair (target fiber i, source fiber j != i) the kernel
// computes the outer integrals
//     Theta^{ij}_{kl} = sum_p w_p P_k(s_p) G_l(x_i + s_p t_i),  k,l = 0..M,
// where G_l is the inner integral over fiber j of the Stokeslet+doublet
// kernel times P_l (3x3 symmetric, 6 components):
//   analytic: fortran/Int_analytical.f90:5-293,663-767 (moments I,J,S with
//             forward recurrence for c<10, backward from n=27 otherwise);
//   numeric : fortran/quadrature.f90:100-164 (LQ-point Gauss quadrature).
// From Theta it writes
//   * A (column-major): rows (i,k,a), k=1..M, cols (j,l,b), l=1..M,
//       k=1: D1 t_a (Theta_kl t)_b ;  k>=2: gamma_k (Theta_kl,ab - E_k t_a (Theta_kl t)_b)
//     (fortran/matrix_assembly.f90:341-395), identity on the diagonal blocks;
//   * partial right-hand sides from Theta_{k0} f0_j
//     (fortran/matrix_assembly.f90:632-905), one slice per source split;
//   * the velocity operator W: rows (i, r in {V,Omega}, c), cols (j,l,c'),
//     W = Theta_{r l} (P_0 and P_1 outer weights of
//     fortran/coeffs_and_vels.f90:612-620,715-724), so the velocities of the
//     step are one GEMV with [F/2, a^1..a^M] per source fiber.
//
// Work decomposition: a workgroup owns T consecutive target fibers and one
// range of source fibers.  Phase 1: thread (t,p) evaluates G_0..G_M at the
// quadrature point p of target t and stores them in LDS.  Phase 2: thread
// (t,l,c) contracts over p into Theta_{0..M,l,c}.  Phase 3: threads own rows
// of A / W and write the 3M (resp. 3(M+1)) columns of source j, so each
// column segment of T*3M consecutive rows is one coalesced store.
#include "rf_inner.h"
#include "rf_internal.h"

#include <algorithm>

namespace rf {
namespace {

struct KArgs {
    int N, n, LQ, T, src_per_split;
    int64_t lda, ldw;
    int analytic, rhs_z_only, write_w;
    double eps, D1, near_c;
    double Ek[kMaxTerms + 1], gam[kMaxTerms + 1];
    const double* pos;
    const double* orient;
    const double* force;
    const double* wp;
    const double* pv;
    void* A;
    void* W;
    double* bpart;
};

constexpr int kMaxItems = 8;  // A-row items per thread (3M/LQ <= 8)

template <int M, bool ANALYTIC>
__host__ __device__ constexpr int chunk_terms()
{
    // Legendre indices l per LDS stage: the analytic path keeps the moments
    // in registers and stages G_l about two l at a time (M = 5: three
    // stages of two; the staging array then takes 19 KB and four
    // workgroups fit a CU's LDS instead of three: 2.24 -> 2.11 ms at
    // N=1000); the numeric path stages all.
    return ANALYTIC ? (M + 3) / 3 : M + 1;
}

template <int M, bool ANALYTIC, bool MIXED>
__global__ void __launch_bounds__(256, 2) assemble_kernel(const KArgs a)
{
    using AT = typename std::conditional<MIXED, float, double>::type;
    constexpr int NL = M + 1;
    constexpr int CH = chunk_terms<M, ANALYTIC>();
    constexpr int NCH = (NL + CH - 1) / CH;
    extern __shared__ double smem[];
    const int T = a.T, LQ = a.LQ, N = a.N;
    const int gstride = LQ + 1;
    double* sG = smem;                                    // [T*CH*6][LQ+1]
    double* sTh = smem + (size_t)T * CH * 6 * gstride;    // [T][NL k][NL l][6]
    double* sW = sTh + (size_t)T * NL * NL * 6;           // [LQ][NL]: w_p P_k(s_p)
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int i0 = blockIdx.x * T;
    const int jb = blockIdx.y * a.src_per_split;
    const int je = min(N, jb + a.src_per_split);
    const double e2 = 2.0 * a.eps * a.eps, e6 = 6.0 * a.eps * a.eps;

    // phase-1 identity: target t1, point p1
    const int t1 = tid / LQ, p1 = tid - (tid / LQ) * LQ;
    const int i1 = i0 + t1;
    const bool on1 = (t1 < T) && (i1 < N);
    double X0 = 0, X1 = 0, X2 = 0;
    if (on1) {
        const double s = a.pv[p1];
        X0 = a.pos[3 * i1 + 0] + s * a.orient[3 * i1 + 0];
        X1 = a.pos[3 * i1 + 1] + s * a.orient[3 * i1 + 1];
        X2 = a.pos[3 * i1 + 2] + s * a.orient[3 * i1 + 2];
    }
    // the Legendre weights of the Theta contraction, once per workgroup
    for (int q = tid; q < LQ * NL; q += nthr) {
        const int p = q / NL, k = q - p * NL;
        sW[q] = a.wp[k * LQ + p];
    }
    // (the first barrier of the source loop orders these stores before the reads)
    const int nA = T * 3 * M;          // A-row items
    const int nItems = nA + (a.write_w ? T * 6 : 0);  // + W-row items
    double bacc[kMaxItems];
#pragma unroll
    for (int s = 0; s < kMaxItems; ++s) bacc[s] = 0.0;

    for (int j = jb; j < je; ++j) {
        const double xj[3] = {a.pos[3 * j], a.pos[3 * j + 1], a.pos[3 * j + 2]};
        const double tj[3] = {a.orient[3 * j], a.orient[3 * j + 1], a.orient[3 * j + 2]};
        const bool act = on1 && (i1 != j);
        // ---- phase 1a: per-point work shared by all l ----
        Moments<M> mo;
        double Gn[ANALYTIC ? 1 : NL * 6];
        if (act) {
            bool fp64 = true;
            if (MIXED) {
                const double dx = X0 - xj[0], dy = X1 - xj[1], dz = X2 - xj[2];
                fp64 = (dx * dx + dy * dy + dz * dz) < a.near_c;
            }
            if constexpr (ANALYTIC) {
                if (fp64)
                    analytic_moments<M, double>(X0, X1, X2, xj, tj, mo);
                else
                    analytic_moments<M, float>(X0, X1, X2, xj, tj, mo);
            } else {
                if (fp64)
                    numeric_G<M, double>(X0, X1, X2, xj, tj, a.eps, LQ, a.pv, a.wp, Gn);
                else
                    numeric_G<M, float>(X0, X1, X2, xj, tj, a.eps, LQ, a.pv, a.wp, Gn);
            }
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            // ---- phase 1b: G_l for the l of this stage into LDS ----
            if (t1 < T) {
#pragma unroll
                for (int li = 0; li < CH; ++li) {
                    const int l = ch * CH + li;
                    if (l <= M) {
                        double g[6] = {0, 0, 0, 0, 0, 0};
                        if (act) {
                            if constexpr (ANALYTIC) {
                                analytic_Gl<M>(mo, l, e2, e6, g);
                            } else {
#pragma unroll
                                for (int q = 0; q < 6; ++q) g[q] = Gn[l * 6 + q];
                            }
                        }
#pragma unroll
                        for (int q = 0; q < 6; ++q) sG[(size_t)((t1 * CH + li) * 6 + q) * gstride + p1] = g[q];
                    }
                }
            }
            __syncthreads();
            // ---- phase 2: Theta_{k l c} = sum_p w_p P_k(s_p) G_{l c}(p) ----
            for (int it = tid; it < T * CH * 6; it += nthr) {
                const int t = it / (CH * 6), rem = it - t * (CH * 6);
                const int li = rem / 6, c = rem - li * 6;
                const int l = ch * CH + li;
                if (l > M) continue;
                const double* gr = sG + (size_t)it * gstride;
                double acc[NL];
#pragma unroll
                for (int k = 0; k < NL; ++k) acc[k] = 0.0;
                for (int p = 0; p < LQ; ++p) {
                    const double gp = gr[p];
                    const double* wk = sW + p * NL;
#pragma unroll
                    for (int k = 0; k < NL; ++k) acc[k] += wk[k] * gp;
                }
#pragma unroll
                for (int k = 0; k < NL; ++k) sTh[((t * NL + k) * NL + l) * 6 + c] = acc[k];
            }
            __syncthreads();
        }
        // ---- phase 3: A rows, W rows, rhs partials ----
        int slot = 0;
        for (int it = tid; it < nItems; it += nthr, ++slot) {
            if (it < nA) {
                const int t = it / (3 * M), rem = it - t * (3 * M);
                const int i = i0 + t;
                if (i >= N) continue;
                const int k = rem / 3 + 1, ar = rem - (rem / 3) * 3;
                const double ti[3] = {a.orient[3 * i], a.orient[3 * i + 1], a.orient[3 * i + 2]};
                const int64_t row = (int64_t)i * 3 * M + rem;
                AT* A = reinterpret_cast<AT*>(a.A);
                if (i == j) {
#pragma unroll 1
                    for (int l = 1; l <= M; ++l)
#pragma unroll
                        for (int bc = 0; bc < 3; ++bc) {
                            const int64_t col = (int64_t)j * 3 * M + (l - 1) * 3 + bc;
                            __builtin_nontemporal_store((AT)((k == l && ar == bc) ? 1.0 : 0.0), &A[col * a.lda + row]);
                        }
                    continue;
                }
                const double* th = sTh + (size_t)(t * NL + k) * NL * 6;
                const double Ek = a.Ek[k], gk = a.gam[k];
#pragma unroll 1
                for (int l = 1; l <= M; ++l) {
                    const double* tl = th + l * 6;
#pragma unroll
                    for (int bc = 0; bc < 3; ++bc) {
                        const double Qb = tl[sym6(bc, 0)] * ti[0] + tl[sym6(bc, 1)] * ti[1] + tl[sym6(bc, 2)] * ti[2];
                        const double v = (k == 1) ? a.D1 * ti[ar] * Qb : gk * (tl[sym6(ar, bc)] - Ek * ti[ar] * Qb);
                        const int64_t col = (int64_t)j * 3 * M + (l - 1) * 3 + bc;
                        __builtin_nontemporal_store((AT)v, &A[col * a.lda + row]);
                    }
                }
                // rhs: contr = Theta_{k0} f0_j
                const double F0 = a.rhs_z_only ? 0.0 : 0.5 * a.force[3 * j + 0];
                const double F1 = a.rhs_z_only ? 0.0 : 0.5 * a.force[3 * j + 1];
                const double F2 = 0.5 * a.force[3 * j + 2];
                const double* t0 = th;  // l = 0
                double cr[3];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
                    cr[cc] = t0[sym6(cc, 0)] * F0 + t0[sym6(cc, 1)] * F1 + t0[sym6(cc, 2)] * F2;
                const double tdc = ti[0] * cr[0] + ti[1] * cr[1] + ti[2] * cr[2];
                const double db = (k == 1) ? a.D1 * tdc * ti[ar] : gk * (cr[ar] - Ek * tdc * ti[ar]);
#pragma unroll
                for (int s = 0; s < kMaxItems; ++s)
                    if (s == slot) bacc[s] -= db;
            } else {
                const int u = it - nA;
                const int t = u / 6, r6 = u - t * 6;
                const int i = i0 + t;
                if (i >= N) continue;
                const int r = r6 / 3, c = r6 - r * 3;
                const int64_t row = (int64_t)i * 6 + r6;
                AT* W = reinterpret_cast<AT*>(a.W);
                const double* th = sTh + (size_t)(t * NL + r) * NL * 6;
#pragma unroll 1
                for (int l = 0; l <= M; ++l)
#pragma unroll
                    for (int c2 = 0; c2 < 3; ++c2) {
                        const int64_t col = (int64_t)j * 3 * NL + l * 3 + c2;
                        __builtin_nontemporal_store((AT)((i == j) ? 0.0 : th[l * 6 + sym6(c, c2)]), &W[col * a.ldw + row]);
                    }
            }
        }
    }
    // rhs partial of this source split
    int slot = 0;
    for (int it = tid; it < nA; it += nthr, ++slot) {
        const int t = it / (3 * M), rem = it - t * (3 * M);
        const int i = i0 + t;
        if (i >= N) continue;
        double v = 0.0;
#pragma unroll
        for (int s = 0; s < kMaxItems; ++s)
            if (s == slot) v = bacc[s];
        a.bpart[(size_t)blockIdx.y * a.n + (size_t)i * 3 * M + rem] = v;
    }
}

template <int M>
hipError_t launch_m(const KArgs& k, dim3 grid, int threads, size_t lds, bool analytic, bool mixed,
                    hipStream_t s)
{
    if (analytic && !mixed)
        hipLaunchKernelGGL((assemble_kernel<M, true, false>), grid, dim3(threads), lds, s, k);
    else if (analytic && mixed)
        hipLaunchKernelGGL((assemble_kernel<M, true, true>), grid, dim3(threads), lds, s, k);
    else if (!analytic && !mixed)
        hipLaunchKernelGGL((assemble_kernel<M, false, false>), grid, dim3(threads), lds, s, k);
    else
        hipLaunchKernelGGL((assemble_kernel<M, false, true>), grid, dim3(threads), lds, s, k);
    return hipGetLastError();
}

}  // namespace

size_t assemble_lds_bytes(int T, int LQ, int M, bool analytic)
{
    const size_t NL = M + 1;
    const size_t CH = analytic ? (M + 3) / 3 : M + 1;  // chunk_terms<M, ANALYTIC>()
    return sizeof(double) * ((size_t)T * CH * 6 * (LQ + 1) + (size_t)T * NL * NL * 6 + (size_t)LQ * NL);
}

int assemble_targets_per_block(int LQ, int M, bool analytic)
{
    // T*LQ threads; prefer a whole number of 64-wide wavefronts <= 256 whose
    // LDS fits in the 160 KiB of a CU.
    const size_t cap = 160 * 1024;
    for (int thr = 256; thr >= 64; thr -= 64)
        if (thr % LQ == 0 && assemble_lds_bytes(thr / LQ, LQ, M, analytic) <= cap) return thr / LQ;
    int T = std::max(1, 256 / LQ);
    while (T > 1 && assemble_lds_bytes(T, LQ, M, analytic) > cap) --T;
    return T;
}

hipError_t launch_assemble(const Tables& t, const AssembleArgs& a, hipStream_t s)
{
    const int M = t.M, LQ = t.LQ;
    const int T = a.targets_per_block > 0 ? a.targets_per_block : assemble_targets_per_block(LQ, M, a.analytic != 0);
    const int threads = T * LQ;
    if (threads > 1024 || T * 3 * M > kMaxItems * threads) return hipErrorInvalidValue;
    KArgs k;
    k.N = a.N;
    k.n = a.n;
    k.LQ = LQ;
    k.T = T;
    k.src_per_split = (a.N + a.splits - 1) / a.splits;
    k.lda = a.lda;
    k.ldw = a.ldw;
    k.analytic = a.analytic;
    k.rhs_z_only = a.analytic ? a.rhs_z_only : 0;
    k.write_w = (a.W != nullptr) ? 1 : 0;
    k.eps = t.eps;
    k.D1 = t.D1;
    k.near_c = a.near_c;
    for (int i = 0; i <= kMaxTerms; ++i) {
        k.Ek[i] = t.Ek[i];
        k.gam[i] = t.gammak[i];
    }
    k.pos = a.pos;
    k.orient = a.orient;
    k.force = a.force;
    k.wp = a.wp;
    k.pv = a.pv;
    k.A = a.A;
    k.W = a.W;
    k.bpart = a.bpart;
    const size_t lds = assemble_lds_bytes(T, LQ, M, a.analytic != 0);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    dim3 grid((a.N + T - 1) / T, a.splits);
    const bool an = a.analytic != 0, mx = a.mixed != 0;
    switch (M) {
    case 1: return launch_m<1>(k, grid, threads, lds, an, mx, s);
    case 2: return launch_m<2>(k, grid, threads, lds, an, mx, s);
    case 3: return launch_m<3>(k, grid, threads, lds, an, mx, s);
    case 4: return launch_m<4>(k, grid, threads, lds, an, mx, s);
    case 5: return launch_m<5>(k, grid, threads, lds, an, mx, s);
    case 6: return launch_m<6>(k, grid, threads, lds, an, mx, s);
    case 7: return launch_m<7>(k, grid, threads, lds, an, mx, s);
    case 8: return launch_m<8>(k, grid, threads, lds, an, mx, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace rf
