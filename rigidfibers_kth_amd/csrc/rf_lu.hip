// rf_lu.hip -- dense LU with partial pivoting on gfx950 (the direct solver,
// LAPACK dgesv at fortran/coeffs_and_vels.f90:62,158,333,514).
//
// Blocked right-looking factorisation, column-major, panel width NB, the
// trailing update delayed to every second panel (super-panels of 2 NB):
//   for each panel [k0, k1) (first bringing its columns up to date):
//     either the block-pivoted panel (diagonal block in LDS, L21 = A21 U11^-1,
//       checked against partial pivoting) or the column-by-column (Crout)
//       panel with the pivot search over the whole column;
//     apply the panel's row swaps to the other columns (laswp);
//     U12 = L11^-1 A12 (unit lower triangular solve, the panel's rows first
//       brought up to date);
//   A22 -= L21 U12 once per super-panel on the matrix cores
//     (v_mfma_f64_16x16x4f64 for FP64, v_mfma_f32_16x16x4f32 for FP32
//     factors, or v_mfma_f32_16x16x32_f16 on fp16-rounded operands for FP32
//     factors that only precondition).
// Pivots: first index of the largest |a_ik| (LAPACK idamax), so the pivot
// sequence is dgesv's whenever the arithmetic agrees.
#include "rf_internal.h"
#include "rf_lu.h"

#include <algorithm>
#include <chrono>
#include <type_traits>
#include <utility>
#include <vector>

namespace rf {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kColThreads = 256;
constexpr int kColMaxBlocks = 512;
constexpr int kMaxPanel = 128;  // widest panel: 128 for FP32 factors (64 by default), 64 for FP64

// (value, index) maximum with ties to the smaller index (LAPACK idamax keeps
// the first maximal entry) over one workgroup; result valid in thread 0.
__device__ __forceinline__ void block_argmax(double& best, int& bi, double* sv, int* si)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) {
            best = ov;
            bi = oi;
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sv[w] = best;
        si[w] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x / 64); ++q)
            if (sv[q] > best || (sv[q] == best && si[q] < bi)) {
                best = sv[q];
                bi = si[q];
            }
    }
}

// One column of the left-looking (Crout) panel factorisation; one launch per
// column j of the panel [k0, k1).  On entry the panel columns < j hold L
// (column j-1 not yet divided by its pivot), rows [k0, j) of column j hold
// U[k0..j, j], and every earlier row swap of the panel has been applied to
// all panel columns.  The launch
//  * divides column j-1 by its pivot (rows >= j),
//  * forms v_i = A[i,j] - sum_{q in [k0,j)} L[i,q] U[q,j] for rows i >= j
//    and stores it in column j,
//  * and the last workgroup to finish (atomic counter) picks the pivot of
//    column j (first largest |v_i|, LAPACK idamax), records ipiv[j], swaps
//    rows j and p over the panel columns, and forms row j of U for the panel
//    columns right of j (one short dot product per column), so that the next
//    launch finds U[k0..j+1, j+1] complete.
// Only two columns are written per launch (the right-looking form rewrites
// the whole rest of the panel every column).
template <typename T>
__global__ void __launch_bounds__(kColThreads) lu_crout_kernel(T* A, int64_t lda, int n, int k0, int k1, int j,
                                                               int* ipiv, int* info, double* pval, int* pidx,
                                                               unsigned int* counter)
{
    __shared__ double sU[kMaxPanel];   // U[k0..j, j]
    __shared__ double sLr[kMaxPanel];  // row j of L (after the swap)
    __shared__ double sv[kColThreads / 64];
    __shared__ int si[kColThreads / 64];
    __shared__ int last_s, piv_s;
    const int m = j - k0;  // columns of L to the left of j inside the panel
    T* cj = A + (int64_t)j * lda;
    for (int q = threadIdx.x; q < m; q += kColThreads) sU[q] = (double)cj[k0 + q];
    double inv = 0.0;
    if (m > 0) {
        const double piv = (double)A[(int64_t)(j - 1) * lda + (j - 1)];
        inv = (piv != 0.0) ? 1.0 / piv : 0.0;  // zero pivot: info is already set
    }
    __syncthreads();
    double best = -1.0;
    int bi = n;
    for (int i = j + blockIdx.x * kColThreads + threadIdx.x; i < n; i += gridDim.x * kColThreads) {
        double v = (double)cj[i];
        if (m > 0) {
            T* cp = A + (int64_t)(j - 1) * lda;
            const double l = (double)cp[i] * inv;
            cp[i] = (T)l;
            const T* li = A + (int64_t)k0 * lda + i;
            int q = 0;
            for (; q + 8 <= m - 1; q += 8) {  // eight loads in flight
                double a[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] = (double)li[(int64_t)(q + u) * lda];
#pragma unroll
                for (int u = 0; u < 8; ++u) v -= a[u] * sU[q + u];
            }
            for (; q < m - 1; ++q) v -= (double)li[(int64_t)q * lda] * sU[q];
            v -= l * sU[m - 1];
        }
        const T vt = (T)v;
        cj[i] = vt;
        const double a = fabs((double)vt);
        if (a > best) {  // strict: first index among equals in this thread
            best = a;
            bi = i;
        }
    }
    block_argmax(best, bi, sv, si);
    if (threadIdx.x == 0) {
        pval[blockIdx.x] = best;
        pidx[blockIdx.x] = bi;
    }
    __syncthreads();  // the workgroup's matrix and partial stores are issued
    if (threadIdx.x == 0) {
        __threadfence();  // ... and made visible device-wide before the count
        last_s = (atomicAdd(counter, 1u) == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!last_s) return;
    __threadfence();
    best = -1.0;
    bi = n;
    for (int q = threadIdx.x; q < (int)gridDim.x; q += kColThreads) {
        const double v = __hip_atomic_load(&pval[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int ii = __hip_atomic_load(&pidx[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v > best || (v == best && ii < bi)) {
            best = v;
            bi = ii;
        }
    }
    block_argmax(best, bi, sv, si);
    if (threadIdx.x == 0) {
        const int p = (bi < n) ? bi : j;
        piv_s = p;
        ipiv[j] = p;
        if (!(best > 0.0) && *info == 0) *info = j + 1;
        atomicExch(counter, 0u);
    }
    __syncthreads();
    const int p = piv_s;
    if (p != j) {
        for (int c = k0 + threadIdx.x; c < k1; c += kColThreads) {
            T* cc = A + (int64_t)c * lda;
            const T t = cc[j];
            cc[j] = cc[p];
            cc[p] = t;
        }
    }
    __syncthreads();
    // Row j of U for the panel columns right of j (rows above j were done by
    // the earlier launches): U[j,c] = A[j,c] - sum_{q in [k0,j)} L[j,q] U[q,c].
    if (j + 1 >= k1) return;
    for (int q = threadIdx.x; q < m; q += kColThreads) sLr[q] = (double)A[(int64_t)(k0 + q) * lda + j];
    __syncthreads();
    for (int c = j + 1 + threadIdx.x; c < k1; c += kColThreads) {
        T* cc = A + (int64_t)c * lda;
        double u = (double)cc[j];
        int q = 0;
        for (; q + 8 <= m; q += 8) {  // eight loads in flight
            double a[8];
#pragma unroll
            for (int w = 0; w < 8; ++w) a[w] = (double)cc[k0 + q + w];
#pragma unroll
            for (int w = 0; w < 8; ++w) u -= sLr[q + w] * a[w];
        }
        for (; q < m; ++q) u -= sLr[q] * (double)cc[k0 + q];
        cc[j] = (T)u;
    }
}

// Column c of L below the diagonal divided by the pivot (the panel's last
// column, which no later Crout launch scales).
template <typename T>
__global__ void lu_scale_kernel(T* A, int64_t lda, int n, int c)
{
    const int i = c + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double piv = (double)A[(int64_t)c * lda + c];
    const double inv = (piv != 0.0) ? 1.0 / piv : 0.0;
    T* cc = A + (int64_t)c * lda;
    cc[i] = (T)((double)cc[i] * inv);
}

// Row swaps ipiv[k0..k1) applied to columns outside [k0, k1) (LAPACK dlaswp).
template <typename T>
__global__ void lu_laswp_kernel(T* A, int64_t lda, int n, int k0, int k1, const int* __restrict__ ipiv)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int outside = n - (k1 - k0);
    if (c >= outside) return;
    if (c >= k0) c += (k1 - k0);
    T* cc = A + (int64_t)c * lda;
    for (int k = k0; k < k1; ++k) {
        const int p = ipiv[k];
        if (p != k) {
            const T t = cc[k];
            cc[k] = cc[p];
            cc[p] = t;
        }
    }
}

// ---- block-pivoted panel ---------------------------------------------------
// Partial pivoting restricted to the panel's diagonal block: the block is
// factorised by one workgroup in LDS, then L21 = A21 U11^-1 for all rows
// below it.  If every multiplier of L21 satisfies |l| <= 1, each block pivot
// was also a first largest entry of its whole column (rows below the block
// come later in LAPACK idamax order), so pivots and factors are those of
// partial pivoting, up to the order of the roundings.  Otherwise (or at a
// zero pivot) info[1] is set, the host restores the panel from the copy the
// kernels took and factorises it column by column.  Most panels of fiber
// systems pass (near contacts make a few fail), and a passing panel costs
// two launches instead of one per column.

// Reads the nb x nb block of A at (k0, k0) (nb <= NB) with up to 16 loads
// in flight per thread and hands element (r, c) to f(r, c, v).  THREADS
// threads: thread t covers row t % NB of the columns t / NB + (THREADS / NB) u,
// so each load instruction reads consecutive rows of one column.
template <typename T, int NB, int THREADS = 256, typename F>
__device__ __forceinline__ void load_block(const T* A, int64_t lda, int k0, int nb, F&& f)
{
    constexpr int H = THREADS / NB, PER = NB / H, BATCH = PER < 16 ? PER : 16;
    static_assert(PER % BATCH == 0, "columns per thread");
    const int r = threadIdx.x % NB, c0 = threadIdx.x / NB;
#pragma unroll
    for (int u0 = 0; u0 < PER; u0 += BATCH) {
        T v[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const int c = c0 + H * (u0 + u);
            v[u] = (r < nb && c < nb) ? A[(int64_t)(k0 + c) * lda + k0 + r] : T(0);
        }
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const int c = c0 + H * (u0 + u);
            if (r < nb && c < nb) f(r, c, v[u]);
        }
    }
}

// Diagonal block [k0, k0+nb) factorised with pivot search inside the block;
// ipiv[k0..k0+nb) as LAPACK (0-based global rows); bak receives the block's
// original values.  One workgroup of kDiagThreads: thread t owns column
// t % NB and the rows r = t / NB (mod kDiagThreads / NB).  Two barriers per column: (A) the pivot
// row swap and the scaling of column j, (B) the rank-1 update, in which the
// owners of column j+1 also find their partial argmax for the next column.
// The update loads 8 rows before it stores any (LDS latency overlapped).
constexpr int kDiagThreads = 1024;
// columns per trailing update (super-panel): 1024 for FP32 factors, 512 for FP64
template <typename T>
constexpr int super_width() { return sizeof(T) == 4 ? 1024 : 512; }

template <typename T, int NB>
__global__ void __launch_bounds__(kDiagThreads) lu_diag_kernel(T* A, int64_t lda, int k0, int nb, int* ipiv,
                                                               int* info, T* __restrict__ bak, int64_t ldb)
{
    constexpr int H = kDiagThreads / NB;  // row phases
    __shared__ T S[NB][NB + 1];  // S[r][c] = A[k0 + r, k0 + c]
    __shared__ double cand_v[H];
    __shared__ int cand_i[H];
    __shared__ T cand_s[H];
    __shared__ T djj_s;
    __shared__ int piv_l[NB];  // pivots and the zero-pivot flag stay in LDS until the end: a global store
    __shared__ int zero_l;     // before a barrier would make every barrier wait for it
    const int t = threadIdx.x, c = t % NB, h = t / NB;
    if (t == 0) zero_l = 0;
    load_block<T, NB, kDiagThreads>(A, lda, k0, nb, [&](int r, int cc, T v) {
        S[r][cc] = v;
        bak[(int64_t)cc * ldb + r] = v;
    });
    __syncthreads();
    if (c == 0) {  // partial argmax of column 0
        double best = -1.0;
        int bi = nb;
        T bs = T(0);
        for (int r = h; r < nb; r += H) {
            const double v = fabs((double)S[r][0]);
            if (v > best) {
                best = v;
                bi = r;
                bs = S[r][0];
            }
        }
        cand_v[h] = best;
        cand_i[h] = bi;
        cand_s[h] = bs;
        if (h == 0) djj_s = S[0][0];
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        // (A) pivot of column j: first largest |a_rj|, r >= j
        double best = cand_v[0];
        int p = cand_i[0];
        T pv = cand_s[0];
#pragma unroll
        for (int q = 1; q < H; ++q)
            if (cand_v[q] > best || (cand_v[q] == best && cand_i[q] < p)) {
                best = cand_v[q];
                p = cand_i[q];
                pv = cand_s[q];
            }
        const T djj = djj_s;
        if (t == 0) {
            piv_l[j] = p;
            if (!(best > 0.0)) zero_l = 1;  // zero pivot: left to the column-by-column panel
            S[j][j] = pv;
        }
        const T inv = (pv != T(0)) ? T(1) / pv : T(0);
        if (p != j && h == 0 && c != j && c < nb) {  // one thread per column
            const T x = S[j][c];
            S[j][c] = S[p][c];
            S[p][c] = x;
        }
        {
            const int r = j + 1 + t;
            if (r < nb) S[r][j] = ((r == p) ? djj : S[r][j]) * inv;
        }
        __syncthreads();
        // (B) trailing update; column j+1's owners keep the partial argmax
        if (c > j && c < nb) {
            const T pj = S[j][c];
            const bool next = (c == j + 1);
            double nbest = -1.0;
            int nbi = nb;
            T nbs = T(0);
            const int r0 = j + 1 + ((h - (j + 1)) % H + H) % H;
            for (int rb = r0; rb < nb; rb += 8 * H) {
                T l[8], v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int r = rb + u * H;
                    if (r < nb) {
                        l[u] = S[r][j];
                        v[u] = S[r][c];
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int r = rb + u * H;
                    if (r < nb) {
                        v[u] -= l[u] * pj;
                        S[r][c] = v[u];
                        if (next) {
                            const double a = fabs((double)v[u]);
                            if (a > nbest) {
                                nbest = a;
                                nbi = r;
                                nbs = v[u];
                            }
                            if (r == j + 1) djj_s = v[u];
                        }
                    }
                }
            }
            if (next) {
                cand_v[h] = nbest;
                cand_i[h] = nbi;
                cand_s[h] = nbs;
            }
        }
        __syncthreads();
    }
    for (int e = t; e < nb * nb; e += kDiagThreads) {
        const int r = e % nb, cc = e / nb;
        A[(int64_t)(k0 + cc) * lda + k0 + r] = S[r][cc];
    }
    if (t < nb) ipiv[k0 + t] = k0 + piv_l[t];
    if (t == 0 && zero_l) info[1] = 1;
}

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N-1>{})
template <typename F, int... J>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, J...>)
{
    (f(std::integral_constant<int, J>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Value of lane `src` (wave-uniform) of a wavefront.
__device__ __forceinline__ float lane_value(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_value(double v, int src)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// max(v, v of the lane selected by DPP control CTRL) for the rows in ROWMASK;
// lanes without a source (or outside ROWMASK) keep v.  v >= -1 here.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_fmax(double v)
{
    const unsigned long long vb = (unsigned long long)__double_as_longlong(v);
    const unsigned long long ib = (unsigned long long)__double_as_longlong(-2.0);
    const unsigned lo =
        (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)ib, (int)(unsigned)vb, CTRL, ROWMASK, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(ib >> 32), (int)(unsigned)(vb >> 32), CTRL,
                                                              ROWMASK, 0xF, false);
    return fmax(v, __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo)));
}

// Largest v over the wavefront: prefix maxima inside each row of 16 lanes
// (row_shr 1, 2, 4, 8), then across rows (row_bcast 15 and 31); lane 63 ends
// with the maximum.
__device__ __forceinline__ double wave_max(double v)
{
    v = dpp_fmax<0x111, 0xF>(v);
    v = dpp_fmax<0x112, 0xF>(v);
    v = dpp_fmax<0x114, 0xF>(v);
    v = dpp_fmax<0x118, 0xF>(v);
    v = dpp_fmax<0x142, 0xA>(v);
    v = dpp_fmax<0x143, 0xC>(v);
    return lane_value(v, 63);
}

// Diagonal block of width nb <= 64 by four wavefronts: lane r of wavefront w
// holds row r of the columns [16 w, 16 w + 16) in registers (the column loop
// fully unrolled).  For column j the wavefront that owns it finds the pivot
// (the first largest |a_rj| over the rows not yet pivots, ties to the smaller
// position in LAPACK's swapped order, as idamax), scales the column and
// publishes the multipliers (0 for pivot rows) and the pivot lane in LDS;
// after one barrier every wavefront subtracts l_r u_c from its columns c > j,
// the pivot row's u_c coming from readlane.  Rows are not moved while the
// block is factorised: each lane keeps the position its row holds in the
// swapped order and every row is stored at its final position at the end.
// Non-pivot rows take exactly the operations of a row-by-row elimination;
// pivot rows take l = 0, which leaves them unchanged up to the sign of an
// exact zero.  bak receives the original values, ipiv and info[1] as in
// lu_diag_kernel.
constexpr int kDiagCols = 16;  // columns per wavefront
template <typename T>
__global__ void __launch_bounds__(256) lu_diag64_kernel(T* A, int64_t lda, int k0, int nb, int* ipiv, int* info,
                                                        T* __restrict__ bak, int64_t ldb)
{
    __shared__ T sl[2][64];
    __shared__ int sp[2];
    const int r = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;  // wavefront-uniform
    const int c0 = w * kDiagCols;
    const bool live = r < nb;
    T a[kDiagCols];
#pragma unroll
    for (int c = 0; c < kDiagCols; ++c)
        a[c] = (live && c0 + c < nb) ? A[(int64_t)(k0 + c0 + c) * lda + k0 + r] : T(0);
#pragma unroll
    for (int c = 0; c < kDiagCols; ++c)
        if (live && c0 + c < nb) bak[(int64_t)(c0 + c) * ldb + r] = a[c];
    int pos = r, my_ipiv = r;
    bool done = !live;
    bool zero = false;
    // compile-time column indices keep a[] in registers
    static_for<64>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int owner = j / kDiagCols, jl = j % kDiagCols, par = j & 1;
        if (j < nb) {  // uniform over the workgroup
            if (w == owner) {
                const double v = done ? -1.0 : fabs((double)a[jl]);
                const double vmax = wave_max(v);
                const unsigned long long cand = __ballot(v == vmax);
                int p;
                if (__popcll(cand) == 1) {
                    p = __ffsll((unsigned long long)cand) - 1;
                } else {  // exact ties (or no candidate left): the smallest position
                    int k = (v == vmax) ? pos : 1 << 20;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) k = min(k, __shfl_xor(k, off, 64));
                    const unsigned long long at = __ballot(v == vmax && pos == k);
                    p = at ? __ffsll((unsigned long long)at) - 1 : 0;
                }
                p = __builtin_amdgcn_readfirstlane(p);
                if (!(vmax > 0.0)) zero = true;
                const T piv = lane_value(a[jl], p);
                const T inv = (piv != T(0)) ? T(1) / piv : T(0);
                const bool pivot_row = done || r == p;
                const T l = a[jl] * inv;
                if (!pivot_row) a[jl] = l;
                sl[par][r] = pivot_row ? T(0) : l;
                if (r == 0) sp[par] = p;
            }
            __syncthreads();
            const int p = __builtin_amdgcn_readfirstlane(sp[par]);
            const T lr = sl[par][r];
            const int ppos = __builtin_amdgcn_readlane(pos, p);
            if (r == j) my_ipiv = ppos;
            if (r != p && pos == j) pos = ppos;  // the row at position j moves to the pivot's position
            if (r == p) {
                pos = j;
                done = true;
            }
            if (w == owner) {
                static_for<kDiagCols>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    if constexpr (c > jl) a[c] -= lr * lane_value(a[c], p);
                });
            } else if (w > owner) {
                static_for<kDiagCols>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    a[c] -= lr * lane_value(a[c], p);
                });
            }
        }
    });
    if (live) {
#pragma unroll
        for (int c = 0; c < kDiagCols; ++c)
            if (c0 + c < nb) A[(int64_t)(k0 + c0 + c) * lda + k0 + pos] = a[c];
        if (w == 0) ipiv[k0 + r] = k0 + my_ipiv;
    }
    if (r == 0 && zero) info[1] = 1;  // zero pivot: left to the column-by-column panel
}

// Broadcast of lane SRC of each group of 4 lanes (DPP quad_perm).
template <int SRC>
__device__ __forceinline__ int quad_bcast_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, SRC * 0x55, 0xF, 0xF, false);
}
template <int SRC>
__device__ __forceinline__ float quad_bcast(float v)
{
    return __int_as_float(quad_bcast_i<SRC>(__float_as_int(v)));
}
template <int SRC>
__device__ __forceinline__ double quad_bcast(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)quad_bcast_i<SRC>((int)(unsigned)b);
    const unsigned hi = (unsigned)quad_bcast_i<SRC>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ T quad_bcast_dyn(T v, int src)
{
    switch (src) {
    case 0: return quad_bcast<0>(v);
    case 1: return quad_bcast<1>(v);
    case 2: return quad_bcast<2>(v);
    default: return quad_bcast<3>(v);
    }
}

// L21 = A21 U11^-1 in place for rows [k0+nb, n) of the panel columns, and
// the check |l| <= 1 (info[1] = 1 otherwise).  Each row is solved by 4
// lanes of one quad, NB/4 columns each in registers; the solved entry x_c
// is broadcast in the quad and every lane subtracts x_c U[c][d] from its
// columns d.  LDS holds U11 strictly above the diagonal (zeros elsewhere,
// so the update needs no masking) and 1/U[c][c]; 64 rows per workgroup.
template <typename T, int NB>
__global__ void __launch_bounds__(256) lu_l21_kernel(T* A, int64_t lda, int n, int k0, int nb, int* info,
                                                     T* __restrict__ bak, int64_t ldb)
{
    constexpr int CW = NB / 4;                  // columns per lane
    constexpr int CP = CW + 16 / (int)sizeof(T); // group stride: 16-byte offsets spread the 4 groups over banks
    __shared__ __attribute__((aligned(16))) T U[NB][4 * CP];
    __shared__ T invd[NB];
    const int k1 = k0 + nb;
    for (int e = threadIdx.x; e < NB * 4 * CP; e += 256) (&U[0][0])[e] = T(0);
    for (int e = threadIdx.x; e < NB; e += 256) invd[e] = T(0);
    __syncthreads();
    load_block<T, NB>(A, lda, k0, nb, [&](int c, int d, T v) {  // U11[c][d]
        if (d > c) U[c][(d / CW) * CP + d % CW] = v;
        if (d == c) invd[c] = T(1) / v;
    });
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane & 3;
    const int row = k1 + blockIdx.x * 64 + wave * 16 + (lane >> 2);
    const bool live = row < n;
    T a[CW];
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const int col = g * CW + k;
        a[k] = (live && col < nb) ? A[(int64_t)(k0 + col) * lda + row] : T(0);
        if (live && col < nb) bak[(int64_t)col * ldb + (row - k0)] = a[k];
    }
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
#pragma unroll
        for (int cl = 0; cl < CW; ++cl) {
            const int c = gg * CW + cl;
            const T xc = quad_bcast_dyn(a[cl] * invd[c], gg);  // x_c, solved by the quad's lane gg
            if (g == gg) a[cl] = xc;
            const T* u = &U[c][g * CP];
#pragma unroll
            for (int k = 0; k < CW; ++k) a[k] -= xc * u[k];
        }
    }
    bool big = false;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const int col = g * CW + k;
        if (live && col < nb) {
            A[(int64_t)(k0 + col) * lda + row] = a[k];
            big |= !(fabs((double)a[k]) <= 1.0);  // NaN included
        }
    }
    if (big) atomicOr(&info[1], 1);
}

// U12 = L11^-1 A12 (L11 unit lower) in place for the columns [k0+nb, n) of
// the panel rows, and the per-column powers of two for the fp16 update
// (colscale may be null): the largest |u| of each column is taken into
// [2^11, 2^12) (combine: the smaller of this and the scale already there,
// for an update that also uses the U rows of an earlier panel).  Each column is solved by the 4 lanes of a quad, NB/4 rows
// each in registers (the mirror of lu_l21_kernel): x_r is final once the
// rows above it are applied; it is broadcast in the quad and every lane
// subtracts L[q][r] x_r from its rows q > r.  64 columns per workgroup.
template <typename T, int NB>
__global__ void __launch_bounds__(256) lu_u12_kernel(T* A, int64_t lda, int n, int k0, int nb,
                                                     float* __restrict__ colscale, int combine)
{
    constexpr int CW = NB / 4;
    constexpr int CP = CW + 16 / (int)sizeof(T);
    constexpr int VL = 16 / (int)sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VL)));
    __shared__ __attribute__((aligned(16))) T Lt[NB][4 * CP];  // Lt[r][g*CP + k] = L11[g*CW + k][r] below the diagonal
    const int k1 = k0 + nb;
    for (int e = threadIdx.x; e < NB * 4 * CP; e += 256) (&Lt[0][0])[e] = T(0);
    __syncthreads();
    load_block<T, NB>(A, lda, k0, nb, [&](int q, int r, T v) {  // L11[q][r]
        if (q > r) Lt[r][(q / CW) * CP + q % CW] = v;
    });
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane & 3;
    const int col = k1 + blockIdx.x * 64 + wave * 16 + (lane >> 2);
    const bool live = col < n;
    T* src = A + (int64_t)col * lda + k0 + g * CW;
    const bool vec = live && (k0 % VL == 0) && (lda % VL == 0) && (g * CW + CW <= nb);
    T a[CW];
    if (vec) {
#pragma unroll
        for (int k = 0; k < CW; k += VL) {
            const vec_t v = *reinterpret_cast<const vec_t*>(src + k);
#pragma unroll
            for (int u = 0; u < VL; ++u) a[k + u] = v[u];
        }
    } else {
#pragma unroll
        for (int k = 0; k < CW; ++k) a[k] = (live && g * CW + k < nb) ? src[k] : T(0);
    }
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
#pragma unroll
        for (int cl = 0; cl < CW; ++cl) {
            const T xr = quad_bcast_dyn(a[cl], gg);  // x_r from the quad's lane gg
            const T* l = &Lt[gg * CW + cl][g * CP];
#pragma unroll
            for (int k = 0; k < CW; ++k) a[k] -= xr * l[k];
        }
    }
    float m = 0.f;
    if (vec) {
#pragma unroll
        for (int k = 0; k < CW; k += VL) {
            vec_t v;
#pragma unroll
            for (int u = 0; u < VL; ++u) {
                v[u] = a[k + u];
                m = fmaxf(m, fabsf((float)a[k + u]));
            }
            *reinterpret_cast<vec_t*>(src + k) = v;
        }
    } else {
#pragma unroll
        for (int k = 0; k < CW; ++k)
            if (live && g * CW + k < nb) {
                src[k] = a[k];
                m = fmaxf(m, fabsf((float)a[k]));
            }
    }
    if (colscale) {
        m = fmaxf(m, __shfl_xor(m, 1, 64));
        m = fmaxf(m, __shfl_xor(m, 2, 64));
        if (g == 0 && live) {
            const int e = (m > 0.f && m <= 3.0e38f) ? ilogbf(m) : 11;
            const float sc = ldexpf(1.f, min(60, max(-60, 11 - e)));
            colscale[col] = combine ? fminf(colscale[col], sc) : sc;
        }
    }
}

// Panel columns [k0, k0+nb), rows [k0, n), back from the copy the two
// kernels above took (a panel whose check failed).
template <typename T>
__global__ void lu_restore_kernel(T* A, int64_t lda, int n, int k0, int nb, const T* __restrict__ bak, int64_t ldb)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (k0 + r < n && c < nb) A[(int64_t)(k0 + c) * lda + k0 + r] = bak[(int64_t)c * ldb + r];
}

// ---- trailing update C -= L21 * U12 ------------------------------------
// Every kernel updates C = A[r0:r1, c0:c1] -= A[r0:r1, ka:kb] A[ka:kb, c0:c1]
// (the L columns and U rows of the panels [ka, kb), all inside A).
struct GemmRange {
    int r0, r1, c0, c1, ka, kb;
};

// VALU version (reference for the matrix-core kernels; any precision):
// workgroup tile 64x64, K staged 16 at a time in LDS.
constexpr int kTile = 64;
constexpr int kKc = 16;

template <typename T>
__device__ __forceinline__ void stage_tiles(const T* A, int64_t lda, const GemmRange& g, int rbase, int cbase, int kc,
                                            int kcount, double (&As)[kTile][kKc + 1], double (&Bs)[kKc][kTile + 1])
{
    // As[r][kk] = L21[rbase + r, kc + kk] ; Bs[kk][c] = U12[kc + kk, cbase + c]
    for (int e = threadIdx.x; e < kTile * kKc; e += 256) {
        const int r = e % kTile, kk = e / kTile;
        const int row = rbase + r, kcol = kc + kk;
        As[r][kk] = (row < g.r1 && kk < kcount) ? (double)A[(int64_t)kcol * lda + row] : 0.0;
    }
    for (int e = threadIdx.x; e < kTile * kKc; e += 256) {
        const int kk = e % kKc, c = e / kKc;
        const int col = cbase + c, krow = kc + kk;
        Bs[kk][c] = (col < g.c1 && kk < kcount) ? (double)A[(int64_t)col * lda + krow] : 0.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) lu_gemm_valu_kernel(T* A, int64_t lda, GemmRange g)
{
    __shared__ double As[kTile][kKc + 1];
    __shared__ double Bs[kKc][kTile + 1];
    const int rbase = g.r0 + blockIdx.x * kTile, cbase = g.c0 + blockIdx.y * kTile;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    double acc[4][4] = {};
    for (int kc = g.ka; kc < g.kb; kc += kKc) {
        const int kcount = min(kKc, g.kb - kc);
        stage_tiles<T>(A, lda, g, rbase, cbase, kc, kcount, As, Bs);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKc; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = As[ty * 4 + u][kk];
                b[u] = Bs[kk][tx * 4 + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = rbase + ty * 4 + u, col = cbase + tx * 4 + v;
            if (row < g.r1 && col < g.c1) {
                T* p = A + (int64_t)col * lda + row;
                *p = (T)((double)*p - acc[u][v]);
            }
        }
}

// Matrix-core version: workgroup tile BM x BN (BM = 64 in FP64; in FP32 128,
// or 64 for the narrow updates inside a super-panel; BN = 128 in FP32, 64 in
// FP64),
// 4 wavefronts of BM/2 x BN/2 (blocks of 16x16), K staged 32 at a time in LDS
// in the matrix's precision.  Every element of C takes the same products in
// the same order whatever BM is.
// For v_mfma_*_16x16x4 lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator r of lane l is D[4(l/16) + r][l%16] for the f32 instruction
// and D[l/16 + 4r][l%16] for the f64 one (rf_selftest_mfma_layout).  The
// f32 product is D = L21 U12, so a lane's 4 values are 4 consecutive rows
// of one column (one 16-byte access); the f64 product is taken transposed,
// D = (L21 U12)^T, so that for each r a wavefront touches 16 consecutive
// rows of 4 columns.
constexpr int kBM = 128, kBMNarrow = 64, kKC = 32;
// LDS rows padded by 16 bytes (keeps the 16-byte staging stores aligned)
template <typename T>
constexpr int lds_pad() { return 16 / (int)sizeof(T); }
// columns per workgroup tile: 128 in FP32; 64 in FP64 keeps the accumulators
// and the prefetch registers of a wavefront within three waves per SIMD (at
// 64-row tiles)
template <typename T>
constexpr int gemm_bn() { return sizeof(T) == 8 ? 64 : 128; }

__device__ __forceinline__ void mfma_16x16x4(double a, double b, d4& acc)
{
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ void mfma_16x16x4(float a, float b, f4& acc)
{
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}

// Staging of one K chunk: As[kk][r] = L21[rbase + r, kc + kk],
// Bs[kk][c] = U12[kc + kk, cbase + c].  With VEC the global reads are 16-byte
// vectors along the contiguous dimension (rows of L21, K of U12) and are
// issued into registers one chunk ahead of the matrix-core work.
template <typename T, int BM>
struct GemmStage {
    static constexpr int VL = 16 / (int)sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VL)));
    static constexpr int NA = kKC * BM / VL / 256;  // vectors per thread
    static constexpr int BN = gemm_bn<T>();
    static constexpr int NB = kKC * BN / VL / 256;
    vec_t a[NA], b[NB];

    __device__ __forceinline__ void load(const T* A, int64_t lda, int r1, int c1, int rbase, int cbase, int kc,
                                         int kcount, bool vec)
    {
#pragma unroll
        for (int t = 0; t < NA; ++t) {
            const int v = threadIdx.x + 256 * t;
            const int kk = v / (BM / VL), r0 = (v % (BM / VL)) * VL;
            const int row = rbase + r0;
            const T* src = A + (int64_t)(kc + kk) * lda + row;
            // FP64: -L21, the accumulators start at C (see the kernel)
            constexpr T sgn = (sizeof(T) == 8) ? T(-1) : T(1);
            if (vec && kk < kcount && row + VL <= r1) {
                a[t] = sgn * *reinterpret_cast<const vec_t*>(src);
            } else {
#pragma unroll
                for (int u = 0; u < VL; ++u) a[t][u] = (kk < kcount && row + u < r1) ? sgn * src[u] : T(0);
            }
        }
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            const int v = threadIdx.x + 256 * t;
            const int c = v / (kKC / VL), kk0 = (v % (kKC / VL)) * VL;
            const int col = cbase + c;
            const T* src = A + (int64_t)col * lda + kc + kk0;
            if (vec && col < c1 && kk0 + VL <= kcount) {
                b[t] = *reinterpret_cast<const vec_t*>(src);
            } else {
#pragma unroll
                for (int u = 0; u < VL; ++u) b[t][u] = (col < c1 && kk0 + u < kcount) ? src[u] : T(0);
            }
        }
    }
    __device__ __forceinline__ void store(T (&As)[kKC][BM + lds_pad<T>()], T (&Bs)[kKC][BN + lds_pad<T>()]) const
    {
#pragma unroll
        for (int t = 0; t < NA; ++t) {
            const int v = threadIdx.x + 256 * t;
            const int kk = v / (BM / VL), r0 = (v % (BM / VL)) * VL;
            *reinterpret_cast<vec_t*>(&As[kk][r0]) = a[t];
        }
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            const int v = threadIdx.x + 256 * t;
            const int c = v / (kKC / VL), kk0 = (v % (kKC / VL)) * VL;
#pragma unroll
            for (int u = 0; u < VL; ++u) Bs[kk0 + u][c] = b[t][u];
        }
    }
};

template <typename T, int BM>
__global__ void __launch_bounds__(256) lu_gemm_mfma_kernel(T* A, int64_t lda, GemmRange g)
{
    using V4 = typename std::conditional<sizeof(T) == 8, d4, f4>::type;
    constexpr bool kF64 = sizeof(T) == 8;
    constexpr int BN = gemm_bn<T>(), JB = BN / 32;  // JB 16-column blocks per wavefront
    constexpr int IB = BM / 32;                     // IB 16-row blocks per wavefront
    __shared__ __attribute__((aligned(16))) T As[kKC][BM + lds_pad<T>()];
    __shared__ __attribute__((aligned(16))) T Bs[kKC][BN + lds_pad<T>()];
    const int rbase = g.r0 + blockIdx.x * BM, cbase = g.c0 + blockIdx.y * BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * (BM / 2), wc = (wave & 1) * (BN / 2);
    const int lr = lane & 15, lk = lane >> 4;
    // FP64: the accumulators start as the C tile (loads issued ahead of the
    // K loop) and the MFMAs add (-L21) U12, hiding the tile's read latency
    // behind the first chunk.  FP32 keeps zero accumulators and subtracts at
    // the end: preloading C there costs a wavefront per SIMD of occupancy.
    V4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            if constexpr (kF64) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + wr + 16 * i + lr;
                    const int col = cbase + wc + 16 * j + lk + 4 * r;
                    acc[i][j][r] = (row < g.r1 && col < g.c1) ? A[(int64_t)col * lda + row] : T(0);
                }
            } else {
                acc[i][j] = V4{0, 0, 0, 0};
            }
        }
    GemmStage<T, BM> st;
    // 16-byte vectors need rows and K offsets that are multiples of the vector
    const bool vec = (g.ka % GemmStage<T, BM>::VL == 0) && (g.r0 % GemmStage<T, BM>::VL == 0) && (lda % GemmStage<T, BM>::VL == 0);
    st.load(A, lda, g.r1, g.c1, rbase, cbase, g.ka, min(kKC, g.kb - g.ka), vec);
    for (int kc = g.ka; kc < g.kb; kc += kKC) {
        __syncthreads();  // the previous chunk's reads of As/Bs are done
        st.store(As, Bs);
        __syncthreads();
        if (kc + kKC < g.kb) st.load(A, lda, g.r1, g.c1, rbase, cbase, kc + kKC, min(kKC, g.kb - kc - kKC), vec);
#pragma unroll
        for (int ks = 0; ks < kKC; ks += 4) {
            T a[IB], b[JB];
#pragma unroll
            for (int i = 0; i < IB; ++i) a[i] = As[ks + lk][wr + 16 * i + lr];
#pragma unroll
            for (int j = 0; j < JB; ++j) b[j] = Bs[ks + lk][wc + 16 * j + lr];
#pragma unroll
            for (int i = 0; i < IB; ++i)
#pragma unroll
                for (int j = 0; j < JB; ++j) {
                    if constexpr (kF64)
                        mfma_16x16x4(b[j], a[i], acc[i][j]);
                    else
                        mfma_16x16x4(a[i], b[j], acc[i][j]);
                }
        }
    }
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            if constexpr (kF64) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + wr + 16 * i + lr;
                    const int col = cbase + wc + 16 * j + lk + 4 * r;
                    if (row < g.r1 && col < g.c1) A[(int64_t)col * lda + row] = acc[i][j][r];
                }
            } else {
                const int row0 = rbase + wr + 16 * i + 4 * lk;
                const int col = cbase + wc + 16 * j + lr;
                if (col < g.c1) {
                    T* p = A + (int64_t)col * lda + row0;
                    if (row0 + 3 < g.r1 && (row0 & 3) == 0 && (lda & 3) == 0) {
                        f4 c = *reinterpret_cast<f4*>(p);
                        c -= acc[i][j];
                        *reinterpret_cast<f4*>(p) = c;
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (row0 + r < g.r1) p[r] -= acc[i][j][r];
                    }
                }
            }
        }
}

// fp16 matrix-core update of FP32 factors (preconditioner factors only):
// C[:, j] = (C[:, j] s_j + (-L21) (U12[:, j] s_j)) / s_j with the operands
// rounded to fp16 and FP32 accumulation (v_mfma_f32_16x16x32_f16).  |L21| <= 1
// under partial pivoting, and s_j (a power of two from the U12 solve) puts
// column j of U12 into [2^11, 2^12), inside fp16's normal range.  Tile TM x
// TN (128 x 128, or 64 x 64 for the narrow updates of a 64-column panel:
// twice the workgroups and no empty half), 4 wavefronts of TM/2 x TN/2
// (16x16 blocks), K staged 32 at a time in LDS
// as fp16 (double-buffered, global loads one chunk ahead); the accumulators
// start as the scaled C tile.  The operand layout follows the 16x16x4 f32
// instruction: lane l supplies row l%16 of A and column l%16 of B at the
// K slots of its group l/16 (8 consecutive K here), and holds
// D[4(l/16)+r][l%16].
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
constexpr int kHB = 128, kHK = 32, kHPad = 8;

template <int TM, int TN>
struct HalfStage {
    static constexpr int RG = TM / 4;        // 4-row groups of A
    static constexpr int AT = RG * kHK / 4;  // threads that load A (4 rows x 4 K each)
    static constexpr int BH = TN * 4 / 256;  // B vectors per thread
    f4 a[4];  // A: rows 4*(t%RG) .. +3 at K columns 4*(t/RG) .. +3 (t < AT)
    f4 b[BH][2];  // B: column v/4, K 8*(v%4) .. +7 for v = t + 256 h
    float bs[BH];

    __device__ __forceinline__ void load(const float* A, int64_t lda, int r1, int c1, int rbase, int cbase, int kc,
                                         int kcount, bool vec, const float* __restrict__ colscale)
    {
        const int t = threadIdx.x;
        const int r0 = rbase + 4 * (t % RG), kk0 = 4 * (t / RG);
#pragma unroll
        for (int q = 0; q < 4 && (AT == 256 || t < AT); ++q) {
            const int kk = kk0 + q;
            const float* src = A + (int64_t)(kc + kk) * lda + r0;
            if (vec && kk < kcount && r0 + 3 < r1) {
                a[q] = *reinterpret_cast<const f4*>(src);
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) a[q][u] = (kk < kcount && r0 + u < r1) ? src[u] : 0.f;
            }
        }
#pragma unroll
        for (int h = 0; h < BH; ++h) {
            const int v = t + 256 * h;
            const int col = cbase + v / 4, kb = 8 * (v % 4);
            const float* src = A + (int64_t)col * lda + kc + kb;
            bs[h] = (col < c1) ? colscale[col] : 0.f;
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                if (vec && col < c1 && kb + 4 * g + 3 < kcount) {
                    b[h][g] = *reinterpret_cast<const f4*>(src + 4 * g);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        b[h][g][u] = (col < c1 && kb + 4 * g + u < kcount) ? src[4 * g + u] : 0.f;
                }
            }
        }
    }
    __device__ __forceinline__ void store(_Float16 (&As)[TM][kHK + kHPad], _Float16 (&Bs)[TN][kHK + kHPad]) const
    {
        const int t = threadIdx.x;
        const int r0 = 4 * (t % RG), kk0 = 4 * (t / RG);
#pragma unroll
        for (int u = 0; u < 4 && (AT == 256 || t < AT); ++u) {
            h4 v;
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = (_Float16)(-a[q][u]);
            *reinterpret_cast<h4*>(&As[r0 + u][kk0]) = v;
        }
#pragma unroll
        for (int h = 0; h < BH; ++h) {
            const int v = t + 256 * h;
            h8 w;
#pragma unroll
            for (int g = 0; g < 2; ++g)
#pragma unroll
                for (int u = 0; u < 4; ++u) w[4 * g + u] = (_Float16)(b[h][g][u] * bs[h]);
            *reinterpret_cast<h8*>(&Bs[v / 4][8 * (v % 4)]) = w;
        }
    }
};

// Staging from the FP32 matrix (conversion and scaling on the way), one K
// chunk of [ka, kb) per call.
template <int TM, int TN>
struct HalfFromMatrix {
    HalfStage<TM, TN> st;
    const float* A;
    int64_t lda;
    int r1, c1, rbase, cbase, ka, kb;
    bool vec;
    const float* colscale;
    __device__ __forceinline__ void load(int chunk)
    {
        const int kc = ka + chunk * kHK;
        st.load(A, lda, r1, c1, rbase, cbase, kc, min(kHK, kb - kc), vec, colscale);
    }
    __device__ __forceinline__ void store(_Float16 (&As)[TM][kHK + kHPad], _Float16 (&Bs)[TN][kHK + kHPad]) const
    {
        st.store(As, Bs);
    }
};

// One TM x TN tile C[rbase.., cbase..] (rows < r1, columns < c1) of the
// fp16 update over nchunks K chunks from the staging object.
template <int TM, int TN, class Stage>
__device__ __forceinline__ void f16_update_tile(float* A, int64_t lda, int rbase, int cbase, int r1, int c1, bool vec,
                                                int nchunks, Stage& st, const float* __restrict__ colscale,
                                                _Float16 (&As)[2][TM][kHK + kHPad],
                                                _Float16 (&Bs)[2][TN][kHK + kHPad])
{
    constexpr int IB = TM / 32, JB = TN / 32;  // 16 x 16 blocks per wavefront
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * (TM / 2), wc = (wave & 1) * (TN / 2);
    const int lr = lane & 15, lg = lane >> 4;
    float sc[JB];
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int col = cbase + wc + 16 * j + lr;
        sc[j] = (col < c1) ? colscale[col] : 1.f;
    }
    f4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            const int row0 = rbase + wr + 16 * i + 4 * lg;
            const int col = cbase + wc + 16 * j + lr;
            const float* p = A + (int64_t)col * lda + row0;
            if (vec && col < c1 && row0 + 3 < r1) {
                acc[i][j] = *reinterpret_cast<const f4*>(p);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[i][j][r] = (col < c1 && row0 + r < r1) ? p[r] : 0.f;
            }
        }
    st.load(0);
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] *= sc[j];
    int buf = 0;
    st.store(As[0], Bs[0]);
    for (int ch = 0; ch < nchunks; ++ch) {
        __syncthreads();  // chunk ch is in As/Bs[buf]; the other buffer's readers are done
        if (ch + 1 < nchunks) st.load(ch + 1);
        h8 a[IB], b[JB];
#pragma unroll
        for (int i = 0; i < IB; ++i) a[i] = *reinterpret_cast<const h8*>(&As[buf][wr + 16 * i + lr][8 * lg]);
#pragma unroll
        for (int j = 0; j < JB; ++j) b[j] = *reinterpret_cast<const h8*>(&Bs[buf][wc + 16 * j + lr][8 * lg]);
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (ch + 1 < nchunks) st.store(As[buf ^ 1], Bs[buf ^ 1]);
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            const int row0 = rbase + wr + 16 * i + 4 * lg;
            const int col = cbase + wc + 16 * j + lr;
            if (col >= c1) continue;
            const f4 c = acc[i][j] * (1.f / sc[j]);
            float* p = A + (int64_t)col * lda + row0;
            if (vec && row0 + 3 < r1) {
                *reinterpret_cast<f4*>(p) = c;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (row0 + r < r1) p[r] = c[r];
            }
        }
}

template <int TM, int TN>
__global__ void __launch_bounds__(256) lu_gemm_f16_kernel(float* A, int64_t lda, GemmRange g,
                                                         const float* __restrict__ colscale)
{
    __shared__ __attribute__((aligned(16))) _Float16 As[2][TM][kHK + kHPad];
    __shared__ __attribute__((aligned(16))) _Float16 Bs[2][TN][kHK + kHPad];
    HalfFromMatrix<TM, TN> st;
    st.A = A;
    st.lda = lda;
    st.r1 = g.r1;
    st.c1 = g.c1;
    st.rbase = g.r0 + blockIdx.x * TM;
    st.cbase = g.c0 + blockIdx.y * TN;
    st.ka = g.ka;
    st.kb = g.kb;
    st.vec = (g.ka % 4 == 0) && (g.r0 % 4 == 0) && (lda % 4 == 0);
    st.colscale = colscale;
    f16_update_tile<TM, TN>(A, lda, st.rbase, st.cbase, g.r1, g.c1, st.vec, (g.kb - g.ka + kHK - 1) / kHK, st,
                            colscale, As, Bs);
}

// ---- packed fp16 operands of a super-panel's trailing update ---------------
// After the panels [K0, K1) of a super-panel: Lp[i][k] = -L[K1 + i][K0 + k]
// and Up[j][k] = U[K0 + k][K1 + j] s_{K1+j} as fp16 with K contiguous (row
// stride kPackK, zero for K1 - K0 <= k < the width rounded to a chunk), so the
// update's staging is one 16-byte copy per row segment.
constexpr int kPackK = super_width<float>();

// L: a 64-row x 64-K tile per workgroup, transposed through LDS.
__global__ void __launch_bounds__(256) pack_l_kernel(const float* __restrict__ A, int64_t lda, int n, int K0, int K1,
                                                     int kwp, _Float16* __restrict__ Lp)
{
    __shared__ float tile[64][65];
    const int t = threadIdx.x;
    const int r0 = K1 + blockIdx.x * 64, kt = blockIdx.y * 64;
    const int kw = K1 - K0;
    {
        const int row = r0 + (t & 63), kq = t >> 6;
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = kt + kq + 4 * u;
            v[u] = (row < n && k < kw) ? A[(int64_t)(K0 + k) * lda + row] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) tile[kq + 4 * u][t & 63] = v[u];
    }
    __syncthreads();
    const int rl = t >> 2, seg = (t & 3) * 16;
    const int row = r0 + rl;
    if (row >= n || kt + seg >= kwp) return;
    h8 w0, w1;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        w0[v] = (_Float16)(-tile[seg + v][rl]);
        w1[v] = (_Float16)(-tile[seg + 8 + v][rl]);
    }
    _Float16* dst = Lp + (size_t)(row - K1) * kPackK + kt + seg;
    *reinterpret_cast<h8*>(dst) = w0;
    *reinterpret_cast<h8*>(dst + 8) = w1;
}

// U: one thread per column and 8-K segment (two 16-byte reads, one write).
__global__ void __launch_bounds__(256) pack_u_kernel(const float* __restrict__ A, int64_t lda, int n, int K0, int K1,
                                                     int kwp, const float* __restrict__ colscale,
                                                     _Float16* __restrict__ Up)
{
    const int segs = kwp / 8;
    const int64_t gidx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int col = K1 + (int)(gidx / segs), seg = (int)(gidx % segs);
    if (col >= n) return;
    const int kw = K1 - K0, k0 = seg * 8;
    const float sc = colscale[col];
    const float* src = A + (int64_t)col * lda + K0 + k0;
    h8 w;
    if (k0 + 8 <= kw && (lda % 4 == 0) && (K0 % 4 == 0)) {
        const f4 x0 = *reinterpret_cast<const f4*>(src), x1 = *reinterpret_cast<const f4*>(src + 4);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            w[v] = (_Float16)(x0[v] * sc);
            w[4 + v] = (_Float16)(x1[v] * sc);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 8; ++v) w[v] = (_Float16)((k0 + v < kw) ? src[v] * sc : 0.f);
    }
    *reinterpret_cast<h8*>(Up + (size_t)(col - K1) * kPackK + k0) = w;
}

// Staging from the packed operands: per chunk each thread copies two
// 16-byte segments of A rows and two of B columns.
struct HalfFromPacked {
    h8 a[2], b[2];
    const _Float16* Lp;
    const _Float16* Up;
    int ri, cj, rn, cn;  // first packed row / column of the tile, packed rows / columns
    __device__ __forceinline__ void load(int chunk)
    {
        const int kc = chunk * kHK;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int v = threadIdx.x + 256 * h;
            const int row = v >> 2, seg = (v & 3) * 8;
            a[h] = (ri + row < rn) ? *reinterpret_cast<const h8*>(Lp + (size_t)(ri + row) * kPackK + kc + seg) : h8{};
            b[h] = (cj + row < cn) ? *reinterpret_cast<const h8*>(Up + (size_t)(cj + row) * kPackK + kc + seg) : h8{};
        }
    }
    __device__ __forceinline__ void store(_Float16 (&As)[kHB][kHK + kHPad], _Float16 (&Bs)[kHB][kHK + kHPad]) const
    {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int v = threadIdx.x + 256 * h;
            *reinterpret_cast<h8*>(&As[v >> 2][(v & 3) * 8]) = a[h];
            *reinterpret_cast<h8*>(&Bs[v >> 2][(v & 3) * 8]) = b[h];
        }
    }
};

// The trailing update of a super-panel from the packed operands: rows and
// columns [K1, n), kwp / 32 chunks of K.  Three workgroups per CU (168
// VGPRs, accumulators in VGPRs): 1.6 % faster than two at n = 60000; four
// spill the tile pointers and are 18 % slower.
__global__ void __launch_bounds__(256, 3) lu_gemm_f16p_kernel(float* A, int64_t lda, int n, int K1, int kwp,
                                                          const _Float16* __restrict__ Lp,
                                                          const _Float16* __restrict__ Up,
                                                          const float* __restrict__ colscale)
{
    __shared__ __attribute__((aligned(16))) _Float16 As[2][kHB][kHK + kHPad];
    __shared__ __attribute__((aligned(16))) _Float16 Bs[2][kHB][kHK + kHPad];
    HalfFromPacked st;
    st.Lp = Lp;
    st.Up = Up;
    st.ri = blockIdx.x * kHB;
    st.cj = blockIdx.y * kHB;
    st.rn = n - K1;
    st.cn = n - K1;
    const bool vec = (K1 % 4 == 0) && (lda % 4 == 0);
    f16_update_tile<kHB, kHB>(A, lda, K1 + st.ri, K1 + st.cj, n, n, vec, kwp / kHK, st, colscale, As, Bs);
}

// The same update with wide tiles: 256 rows, and each wavefront a 128-row
// quadrant whose 8 x (TN/32) blocks of 16 x 16 keep their FP32 accumulators
// in registers for the whole K loop: 43 (TN = 128) or 64 (TN = 256) flops
// per LDS byte read, against 32 for 64 x 64 quadrants.  The K chunks travel
// from the packed operands straight into a ring of LDS stages
// (global_load_lds_dwordx4: no staging registers), issued STAGES - 1 chunks
// ahead.  Stage rows are 64 bytes with no padding; the 16-byte segment s of
// row r sits at s ^ f(r), f(r) = ((r >> 1) ^ (r >> 3)) & 3, so the fragment
// reads (16 rows, one segment) touch every bank once per 8- or 16-lane
// group; each load instruction fills 1 KB of a stage contiguously (16 rows),
// its lanes fetching the segments that the swizzle puts in their slots.
// Measured on the first update of n = 60000 (59000^2 x 1024): with 256 x 256
// tiles, one workgroup per CU, the K loop alone runs at 1.8 PFLOP/s (73 % of
// the dense fp16 peak) but the FP32 C tile's read-modify-write (8 bytes per
// element and super-panel, 28 GB here) does not overlap it: 12.6 ms.  With
// 256 x 128 tiles two workgroups share a CU and one's C traffic overlaps the
// other's products: 12.1 ms, and 240 ms against 270 ms over a factorisation.
constexpr int kWB = 256;  // tile rows
constexpr int kXcds = 8, kCusPerXcd = 32;
// below this trailing size the 128 x 128 tiles fill the chip better
constexpr int kWideMin = 4096;
__device__ __forceinline__ int wide_seg(int r, int s) { return (s ^ ((r >> 1) ^ (r >> 3))) & 3; }

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// 16 bytes per lane from src straight into LDS at dst (wavefront-uniform; lane
// l's bytes land at dst + 16 l).  Kept out of the kernel templates: the host
// pass cannot form the LDS address-space pointer.
__device__ __forceinline__ void ring_load(const _Float16* src, _Float16* dst)
{
    __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)dst, 16, 0, 0);
}

// Barrier without the memory fence of __syncthreads(): the stage data
// arrives by the loads' own counter (waited for explicitly before it), and
// the clobber keeps the compiler from moving LDS reads across it.
__device__ __forceinline__ void stage_barrier() { asm volatile("s_barrier" ::: "memory"); }

// s_waitcnt immediate for "at most v vector-memory operations in flight":
// vmcnt in bits 3:0 and 15:14, expcnt in 6:4 (7: no wait), lgkmcnt in 11:8
// (15: no wait).
constexpr int vm_wait(int v) { return (v & 15) | ((v >> 4) << 14) | 0x0F70; }

// Configuration: TN tile columns (each wavefront 128 x TN/2), STAGES ring
// stages, WG workgroups per CU.  The one used, 256 x 128 with 3 stages (72 KB
// of LDS, 128 accumulators per lane), runs two workgroups per CU, so one
// workgroup's C traffic overlaps the other's products.
template <int TN, int STAGES, int WG>
struct WideCfg {
    static constexpr int kBHalves = TN * kHK;               // U operand of one chunk
    static constexpr int kStageHalves = kWB * kHK + kBHalves;
    static constexpr int kJB = TN / 32;                      // 16-column blocks per wavefront
    static constexpr int kLoads = 4 + TN / 64;               // ring loads per wavefront and chunk
    static constexpr int kBlockRows = (TN == 256) ? 4 : 8;   // tiles per XCD round: WG per CU of the XCD
    static constexpr int kBlockCols = kCusPerXcd * WG / kBlockRows;
};
using WideTwo = WideCfg<128, 3, 2>;

template <int TN, int STAGES, int WG>
__global__ void __launch_bounds__(256, WG) lu_gemm_f16w_kernel(float* A, int64_t lda, int n, int K1, int kwp,
                                                           const _Float16* __restrict__ Lp,
                                                           const _Float16* __restrict__ Up,
                                                           const float* __restrict__ colscale)
{
    using C = WideCfg<TN, STAGES, WG>;
    constexpr int JB = C::kJB;
    __shared__ __attribute__((aligned(1024))) _Float16 ring[STAGES][C::kStageHalves];
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wr = (wave >> 1) * 128, wc = (wave & 1) * (TN / 2);
    const int lr = lane & 15, lg = lane >> 4;
    const int rn = n - K1;
    // Tile of this workgroup: workgroup ids go round-robin over the 8 XCDs,
    // and the workgroups one XCD runs together take a block of tiles, so its
    // L2 serves each chunk of a few L and U panels to all of them (a
    // row-major order would stream a different L panel per workgroup).  Ids
    // past the last tile leave before the first barrier.
    const int TR = (rn + kWB - 1) / kWB, TC = (rn + TN - 1) / TN;
    const int nbi = (TR + C::kBlockRows - 1) / C::kBlockRows, nbj = (TC + C::kBlockCols - 1) / C::kBlockCols;
    constexpr int per = C::kBlockRows * C::kBlockCols;
    const int id = blockIdx.x, xcd = id % kXcds, local = id / kXcds;
    const int blk = (local / per) * kXcds + xcd, slot = local % per;
    const int tx = (blk % nbi) * C::kBlockRows + slot % C::kBlockRows;
    const int ty = (blk / nbi) * C::kBlockCols + slot / C::kBlockRows;
    if (blk >= nbi * nbj || tx >= TR || ty >= TC) return;
    const int ri = tx * kWB, cj = ty * TN;  // packed row / column of the tile
    const int rbase = K1 + ri, cbase = K1 + cj;
    const bool vec = (K1 % 4 == 0) && (lda % 4 == 0);
    const int nchunks = kwp / kHK;
    // this lane's source rows for the wavefront's 16-row groups of each
    // operand (rows past the trailing matrix are clamped to its last row:
    // their products land only in rows / columns that are never stored)
    const _Float16* asrc[4];
    const _Float16* bsrc[TN / 64];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = wave * 64 + q * 16 + (lane >> 2);
        asrc[q] = Lp + (size_t)min(ri + r, rn - 1) * kPackK + wide_seg(r, lane & 3) * 8;
    }
#pragma unroll
    for (int q = 0; q < TN / 64; ++q) {
        const int r = wave * (TN / 4) + q * 16 + (lane >> 2);
        bsrc[q] = Up + (size_t)min(cj + r, rn - 1) * kPackK + wide_seg(r, lane & 3) * 8;
    }
    auto issue = [&](int chunk, int st) {
        const int kc = chunk * kHK;
#pragma unroll
        for (int q = 0; q < 4; ++q) ring_load(asrc[q] + kc, &ring[st][(wave * 64 + q * 16) * kHK]);
#pragma unroll
        for (int q = 0; q < TN / 64; ++q) ring_load(bsrc[q] + kc, &ring[st][kWB * kHK + (wave * (TN / 4) + q * 16) * kHK]);
    };
    // Every iteration issues one chunk, STAGES - 1 ahead of the one it reads;
    // past the last chunk the last one is loaded again into the stage just
    // read (no later iteration reads it), so exactly STAGES - 2 chunks' loads
    // follow the one waited for and the loop has no branches.
    for (int c = 0; c < STAGES - 1; ++c) issue(min(c, nchunks - 1), c);
    f4 acc[8][JB];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    // this lane's 16 bytes of rows r0..r0+15 of operand o (0: L, 1: U) in a stage
    auto frag = [&](int st, int o, int r0) {
        const int r = r0 + lr;
        return *reinterpret_cast<const h8*>(&ring[st][o * kWB * kHK + r * kHK + wide_seg(r, lg) * 8]);
    };
    const bool interior = vec && rbase + kWB <= n && cbase + TN <= n;
    for (int ch = 0; ch < nchunks; ++ch) {
        __builtin_amdgcn_s_waitcnt(vm_wait((STAGES - 2) * C::kLoads));  // this wavefront's chunk ch has landed
        stage_barrier();  // ... and every wavefront's; stage (ch - 1) % STAGES has no readers left
        const int st = ch % STAGES;
        issue(min(ch + STAGES - 1, nchunks - 1), (ch + STAGES - 1) % STAGES);
        h8 a[8], b[JB];
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = frag(st, 0, wr + 16 * i);
#pragma unroll
        for (int j = 0; j < JB; ++j) b[j] = frag(st, 1, wc + 16 * j);
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], b[j], acc[i][j], 0, 0, 0);
        // the next loads as early as possible, the fragment reads, the products
        __builtin_amdgcn_sched_group_barrier(0x020, C::kLoads, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 8 + JB, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 8 * JB, 0);
    }
    __builtin_amdgcn_s_waitcnt(vm_wait(0));  // no load into LDS outlives the workgroup
    // C += (-L)(U s) / s.  Interior tiles: no per-element conditions, all of
    // the wavefront's C loads in flight at once; edge tiles are guarded.
    float inv[JB];
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int col = cbase + wc + 16 * j + lr;
        inv[j] = (col < n) ? 1.f / colscale[col] : 0.f;
    }
    if (interior) {
        f4 c[JB][8];
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
            for (int i = 0; i < 8; ++i)
                c[j][i] = *reinterpret_cast<const f4*>(A + (int64_t)(cbase + wc + 16 * j + lr) * lda + rbase + wr +
                                                       16 * i + 4 * lg);
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
#pragma unroll
                for (int r = 0; r < 4; ++r) c[j][i][r] = fmaf(acc[i][j][r], inv[j], c[j][i][r]);
                *reinterpret_cast<f4*>(A + (int64_t)(cbase + wc + 16 * j + lr) * lda + rbase + wr + 16 * i + 4 * lg) =
                    c[j][i];
            }
        return;
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int col = cbase + wc + 16 * j + lr;
        if (col >= n) continue;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row0 = rbase + wr + 16 * i + 4 * lg;
            float* p = A + (int64_t)col * lda + row0;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (row0 + r < n) p[r] = fmaf(acc[i][j][r], inv[j], p[r]);
        }
    }
}

template <class Cfg, int TN, int STAGES, int WG>
hipError_t launch_wide(float* A, int64_t lda, int n, int K1, int kwp, const _Float16* lpack, const _Float16* upack,
                       const float* colscale, hipStream_t s)
{
    const int rem = n - K1;
    const int TR = (rem + kWB - 1) / kWB, TC = (rem + TN - 1) / TN;
    const int blocks = ((TR + Cfg::kBlockRows - 1) / Cfg::kBlockRows) * ((TC + Cfg::kBlockCols - 1) / Cfg::kBlockCols);
    const int groups = (blocks + kXcds - 1) / kXcds * kXcds * Cfg::kBlockRows * Cfg::kBlockCols;
    hipLaunchKernelGGL((lu_gemm_f16w_kernel<TN, STAGES, WG>), dim3(groups), dim3(256), 0, s, A, lda, n, K1, kwp, lpack,
                       upack, colscale);
    return hipGetLastError();
}

// Block-pivoted panel [k0, k1): diagonal block, then L21 below it; bak
// receives the panel's original columns (rows [k0, n), leading dimension n).
template <typename T>
void block_panel(T* A, int64_t lda, int n, int k0, int k1, int* ipiv, int* info, T* bak, hipStream_t s)
{
    const int nb = k1 - k0;
    const dim3 grid((n - k1 + 63) / 64);
    if constexpr (sizeof(T) == 4) {
        if (nb > 64) {
            hipLaunchKernelGGL((lu_diag_kernel<T, kMaxPanel>), dim3(1), dim3(kDiagThreads), 0, s, A, lda, k0, nb,
                               ipiv, info, bak, (int64_t)n);
            if (k1 < n)
                hipLaunchKernelGGL((lu_l21_kernel<T, kMaxPanel>), grid, dim3(256), 0, s, A, lda, n, k0, nb, info, bak,
                                   (int64_t)n);
            return;
        }
    }
    // panels of at most 64 columns (all FP64 ones): four wavefronts
    hipLaunchKernelGGL(lu_diag64_kernel<T>, dim3(1), dim3(256), 0, s, A, lda, k0, nb, ipiv, info, bak, (int64_t)n);
    if (k1 < n)
        hipLaunchKernelGGL((lu_l21_kernel<T, 64>), grid, dim3(256), 0, s, A, lda, n, k0, nb, info, bak, (int64_t)n);
}

constexpr size_t kWsScaleOffset = kColMaxBlocks * (sizeof(double) + sizeof(int)) + 64;
size_t ws_backup_offset(int n) { return kWsScaleOffset + (sizeof(float) * (size_t)n + 255) / 256 * 256; }
size_t ws_pack_offset(int n) { return ws_backup_offset(n) + (size_t)n * kMaxPanel * sizeof(float); }

// Until *word == seq (a copy kernel stores it after its data at system
// scope): spins up to a second, then waits for the stream and looks again.
static hipError_t wait_host_word(const unsigned* word, unsigned seq, hipStream_t s)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        for (int spin = 0; spin < 4096; ++spin)
            if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == seq) return hipSuccess;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) break;
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    return (__atomic_load_n(word, __ATOMIC_ACQUIRE) == seq) ? hipSuccess : hipErrorUnknown;
}

template <typename T>
hipError_t factor_impl(T* A, int64_t lda, int n, int nb, int* ipiv, int* info, int use_mfma, int block_pivot,
                       int update_fp16, void* ws, int* redone, int* host_flag, hipStream_t s)
{
    double* pval = static_cast<double*>(ws);
    int* pidx = reinterpret_cast<int*>(pval + kColMaxBlocks);
    unsigned int* counter = reinterpret_cast<unsigned int*>(pidx + kColMaxBlocks);
    const bool f16 = (sizeof(T) == 4) && use_mfma && update_fp16;
    float* colscale = f16 ? reinterpret_cast<float*>(static_cast<char*>(ws) + kWsScaleOffset) : nullptr;
    T* bak = reinterpret_cast<T*>(static_cast<char*>(ws) + ws_backup_offset(n));
    _Float16* lpack = reinterpret_cast<_Float16*>(static_cast<char*>(ws) + ws_pack_offset(n));
    _Float16* upack = lpack + (size_t)n * kPackK;
    if (redone) *redone = 0;
    auto gemm = [&](GemmRange g) {
        if (g.r1 <= g.r0 || g.c1 <= g.c0 || g.kb <= g.ka) return;
        const int rows = g.r1 - g.r0, cols = g.c1 - g.c0;
        if (f16) {
            if constexpr (sizeof(T) == 4) {
                // the narrow updates of a panel of at most 64 columns or rows
                // take 64 x 64 tiles (same products per element)
                if (rows <= 64 || cols <= 64)
                    hipLaunchKernelGGL((lu_gemm_f16_kernel<64, 64>), dim3((rows + 63) / 64, (cols + 63) / 64),
                                       dim3(256), 0, s, A, lda, g, colscale);
                else
                    hipLaunchKernelGGL((lu_gemm_f16_kernel<kHB, kHB>), dim3((rows + kHB - 1) / kHB, (cols + kHB - 1) / kHB),
                                       dim3(256), 0, s, A, lda, g, colscale);
            }
        } else if (use_mfma) {
            const int bn = gemm_bn<T>();
            // FP64 takes 64-row tiles everywhere (three workgroups per CU
            // instead of two: 85.7 against 91 ms for n = 15000); FP32 the
            // narrow updates (a panel's columns or rows): twice the
            // workgroups for the columns, no empty half for the rows
            if (sizeof(T) == 8 || rows <= kBMNarrow || cols <= bn)
                hipLaunchKernelGGL((lu_gemm_mfma_kernel<T, kBMNarrow>),
                                   dim3((rows + kBMNarrow - 1) / kBMNarrow, (cols + bn - 1) / bn), dim3(256), 0, s, A,
                                   lda, g);
            else
                hipLaunchKernelGGL((lu_gemm_mfma_kernel<T, kBM>), dim3((rows + kBM - 1) / kBM, (cols + bn - 1) / bn),
                                   dim3(256), 0, s, A, lda, g);
        } else {
            hipLaunchKernelGGL(lu_gemm_valu_kernel<T>, dim3((rows + kTile - 1) / kTile, (cols + kTile - 1) / kTile),
                               dim3(256), 0, s, A, lda, g);
        }
    };
    // Super-panels of two panels: inside one, each panel's columns and rows
    // catch up with the earlier panel (narrow updates), and the trailing
    // matrix takes one update with K = 2 nb per super-panel, which halves its
    // read-modify-write traffic.
    const int sb = std::max(1, super_width<T>() / nb) * nb;
    for (int K0 = 0; K0 < n; K0 += sb) {
        const int K1 = std::min(n, K0 + sb);
        for (int k0 = K0; k0 < K1; k0 += nb) {
            const int k1 = std::min(K1, k0 + nb);
            gemm({k0, n, k0, k1, K0, k0});  // the panel's columns
            bool columns = !block_pivot;
            if (block_pivot) {
                block_panel<T>(A, lda, n, k0, k1, ipiv, info, bak, s);
                // the check decides before anything else reads the panel
                int flag = 0;
                hipError_t e;
                if (host_flag) {  // own copy kernel: the runtime's device-to-host copy stalls the queue
                    // it ends with a sequence word in host_flag[1]: spin on that
                    // rather than sleep in a stream synchronisation
                    unsigned* seq_word = reinterpret_cast<unsigned*>(host_flag + 1);
                    const unsigned seq = __atomic_load_n(seq_word, __ATOMIC_ACQUIRE) + 1u;
                    e = launch_to_host(info + 1, host_flag, sizeof(int), s, seq_word, seq);
                    if (e == hipSuccess) e = wait_host_word(seq_word, seq, s);
                    flag = __atomic_load_n(host_flag, __ATOMIC_ACQUIRE);
                } else {
                    e = hipMemcpyAsync(&flag, info + 1, sizeof(int), hipMemcpyDeviceToHost, s);
                    if (e == hipSuccess) e = hipStreamSynchronize(s);
                }
                if (e != hipSuccess) return e;
                if (flag) {
                    hipLaunchKernelGGL(lu_restore_kernel<T>, dim3((n - k0 + 255) / 256, k1 - k0), dim3(256), 0, s, A,
                                       lda, n, k0, k1 - k0, bak, (int64_t)n);
                    e = hipMemsetAsync(info + 1, 0, sizeof(int), s);
                    if (e != hipSuccess) return e;
                    columns = true;
                    if (redone) *redone += 1;
                }
            }
            for (int j = k0; j < k1 && columns; ++j) {
                const int rows = n - j;
                const int blocks = std::max(1, std::min(kColMaxBlocks, (rows + kColThreads - 1) / kColThreads));
                hipLaunchKernelGGL(lu_crout_kernel<T>, dim3(blocks), dim3(kColThreads), 0, s, A, lda, n, k0, k1, j,
                                   ipiv, info, pval, pidx, counter);
            }
            if (k1 < n && columns)
                hipLaunchKernelGGL(lu_scale_kernel<T>, dim3((n - k1 + 255) / 256), dim3(256), 0, s, A, lda, n, k1 - 1);
            const int outside = n - (k1 - k0);
            if (outside > 0)
                hipLaunchKernelGGL(lu_laswp_kernel<T>, dim3((outside + 255) / 256), dim3(256), 0, s, A, lda, n, k0,
                                   k1, ipiv);
            if (k1 < n) {
                gemm({k0, k1, k1, n, K0, k0});  // the panel's rows right of it
                const dim3 ugrid((n - k1 + 63) / 64);
                const int combine = (k0 > K0) ? 1 : 0;
                bool wide = false;
                if constexpr (sizeof(T) == 4) {
                    if (nb > 64) {
                        hipLaunchKernelGGL((lu_u12_kernel<T, kMaxPanel>), ugrid, dim3(256), 0, s, A, lda, n, k0,
                                           k1 - k0, colscale, combine);
                        wide = true;
                    }
                }
                if (!wide)
                    hipLaunchKernelGGL((lu_u12_kernel<T, 64>), ugrid, dim3(256), 0, s, A, lda, n, k0, k1 - k0, colscale,
                                       combine);
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        if (f16 && K1 < n) {  // the trailing matrix from packed fp16 operands
            if constexpr (sizeof(T) == 4) {
                const int rem = n - K1, kwp = (K1 - K0 + kHK - 1) / kHK * kHK;
                hipLaunchKernelGGL(pack_l_kernel, dim3((rem + 63) / 64, (kwp + 63) / 64), dim3(256), 0, s, A, lda, n,
                                   K0, K1, kwp, lpack);
                const int64_t units = (int64_t)rem * (kwp / 8);
                hipLaunchKernelGGL(pack_u_kernel, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, s, A, lda, n,
                                   K0, K1, kwp, colscale, upack);
                if (rem >= kWideMin) {  // 256 x 128 tiles, two workgroups per CU, in XCD-local blocks
                    const hipError_t e = launch_wide<WideTwo, 128, 3, 2>(A, lda, n, K1, kwp, lpack, upack, colscale, s);
                    if (e != hipSuccess) return e;
                } else {
                    const int tiles = (rem + kHB - 1) / kHB;
                    hipLaunchKernelGGL(lu_gemm_f16p_kernel, dim3(tiles, tiles), dim3(256), 0, s, A, lda, n, K1, kwp,
                                       lpack, upack, colscale);
                }
            }
        } else {
            gemm({K1, n, K1, n, K0, K1});  // the trailing matrix
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// One wavefront: operands encode (mode 0) A[i][0] = i+1, B[0][j] = 1 -> D = i+1;
// (mode 1) A[i][0] = 1, B[0][j] = j+1 -> D = j+1; (mode 2) only the lanes
// 32..47 of A and B are 1 -> D = 1 where both use the same K slot.
__global__ void mfma_probe_kernel(int mode, double* out64, float* out32)
{
    const int lane = threadIdx.x;
    const int grp = lane >> 4, idx = lane & 15;
    double a = 0.0, b = 0.0;
    if (mode == 0) {
        a = (grp == 0) ? idx + 1.0 : 0.0;
        b = (grp == 0) ? 1.0 : 0.0;
    } else if (mode == 1) {
        a = (grp == 0) ? 1.0 : 0.0;
        b = (grp == 0) ? idx + 1.0 : 0.0;
    } else {
        a = (grp == 2) ? 1.0 : 0.0;
        b = (grp == 2) ? 1.0 : 0.0;
    }
    d4 acc = d4{0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    f4 accf = f4{0.f, 0.f, 0.f, 0.f};
    accf = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, accf, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        out64[lane * 4 + r] = acc[r];
        out32[lane * 4 + r] = accf[r];
    }
}

// ---- solves with the factors: permutation + one launch per block -----------
// A solve with preconditioner factors runs once per GMRES iteration, so it
// is built for few launches and no serial chains: once per
// factorisation the pivots become one permutation and the 64x64 diagonal
// blocks of L and U are inverted (one workgroup per block); each 64-row
// block of the forward and backward sweeps is then one launch, in which
// every workgroup forms y = inv(D) v_block (4 lanes per row) and updates 64
// of the remaining rows (each wavefront takes 16 of the block's columns with
// all 16 loads in flight; the 4 partial sums meet in LDS).
constexpr int kSB = 64;

__global__ void permute_kernel(const double* __restrict__ x, const int* __restrict__ perm, int n,
                               double* __restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[perm[i]];
}

// inv[(2 b + u) 64^2 + 64 r + q] = (D_b^-1)[r][q], D_b the diagonal block b of
// L (u = 0, unit lower) or U (u = 1).  Thread c computes column c of the
// inverse by substitution and touches only that column: no barriers.
template <typename T>
__global__ void __launch_bounds__(kSB) block_inverse_kernel(const T* __restrict__ LU, int64_t lda, int n,
                                                            T* __restrict__ inv)
{
    __shared__ T D[kSB][kSB + 1];
    __shared__ T X[kSB][kSB + 1];
    const int b = blockIdx.x, upper = blockIdx.y;
    const int k0 = b * kSB, bs = min(kSB, n - k0);
    const int c = threadIdx.x;
#pragma unroll
    for (int q0 = 0; q0 < kSB; q0 += 16) {
        T v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = q0 + u;
            v[u] = (c < bs && q < bs) ? LU[(int64_t)(k0 + q) * lda + k0 + c] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) D[c][q0 + u] = v[u];
    }
    __syncthreads();
    if (!upper) {
        for (int r = 0; r < kSB; ++r) {  // X[r][c] = delta_rc - sum_{c <= q < r} L[r][q] X[q][c]
            T acc = (r == c) ? T(1) : T(0);
            for (int q = c; q < r; ++q) acc -= D[r][q] * X[q][c];
            X[r][c] = (r < bs && c < bs) ? acc : T(0);
        }
    } else {
        for (int r = kSB - 1; r >= 0; --r) {  // X[r][c] = (delta_rc - sum_{r < q <= c} U[r][q] X[q][c]) / U[r][r]
            T acc = (r == c) ? T(1) : T(0);
            for (int q = r + 1; q <= c; ++q) acc -= D[r][q] * X[q][c];
            X[r][c] = (r < bs && c < bs && r <= c) ? acc / D[r][r] : T(0);
        }
    }
    __syncthreads();
    T* out = inv + ((size_t)b * 2 + upper) * kSB * kSB;
    for (int r = 0; r < kSB; ++r) out[r * kSB + c] = X[r][c];
}

// lower = 1: forward block (unit lower L), rows updated i >= k0+bs;
// lower = 0: backward block (U), rows updated i < k0.
template <typename T, int LOWER>
__global__ void __launch_bounds__(256) sweep_block_kernel(const T* __restrict__ LU, int64_t lda, int n,
                                                          const T* __restrict__ inv, double* __restrict__ v,
                                                          double* __restrict__ out, int k0, int bs)
{
    constexpr int QW = kSB / 4;  // inverse columns per lane, block columns per wavefront
    __shared__ double vb[kSB];
    __shared__ double yk[kSB];
    __shared__ double part[4][64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < kSB) vb[t] = (t < bs) ? v[k0 + t] : 0.0;
    const T* X = inv + ((size_t)(k0 / kSB) * 2 + (LOWER ? 0 : 1)) * kSB * kSB;
    const int r = t >> 2, qb = (t & 3) * QW;
    T xr[QW];
#pragma unroll
    for (int u = 0; u < QW; ++u) xr[u] = X[r * kSB + qb + u];
    const int i = (LOWER ? k0 + bs : 0) + blockIdx.x * 64 + lane;
    const int rend = LOWER ? n : k0;
    const int q0 = wave * QW;
    T a[QW];
    if (i < rend) {
        const T* li = LU + (int64_t)(k0 + q0) * lda + i;
#pragma unroll
        for (int u = 0; u < QW; ++u) a[u] = (q0 + u < bs) ? li[(int64_t)u * lda] : T(0);
    }
    __syncthreads();
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < QW; ++u) sum += (double)xr[u] * vb[qb + u];
    sum += __shfl_xor(sum, 1, 64);
    sum += __shfl_xor(sum, 2, 64);
    if ((t & 3) == 0) yk[r] = sum;
    __syncthreads();
    if (blockIdx.x == 0 && t < bs) out[k0 + t] = yk[t];
    double acc = 0.0;
    if (i < rend) {
#pragma unroll
        for (int u = 0; u < QW; ++u) acc += (double)a[u] * yk[q0 + u];
    }
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && i < rend) v[i] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

}  // namespace

hipError_t mfma_probe(int mode, double* out64, float* out32, hipStream_t s)
{
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, mode, out64, out32);
    return hipGetLastError();
}

// counter and pivot partials, the fp16 column scales, the panel copy (n rows
// x 128 FP32 or 64 FP64 columns)
size_t lu_workspace_bytes(int n, int update_fp16)
{
    return ws_pack_offset(n) + (update_fp16 ? 2 * (size_t)n * kPackK * sizeof(_Float16) : 0);
}

hipError_t lu_factor(void* A, int elem_bytes, int64_t lda, int n, int nb, int* ipiv, int* info, int use_mfma,
                     int block_pivot, int update_fp16, void* ws, int* redone, int* host_flag, hipStream_t s)
{
    if (nb < 1 || nb > (elem_bytes == 4 ? kMaxPanel : 64) || !ws || lda % 4 != 0) return hipErrorInvalidValue;
    if (elem_bytes == 8)
        return factor_impl<double>(static_cast<double*>(A), lda, n, nb, ipiv, info, use_mfma, block_pivot, 0, ws,
                                   redone, host_flag, s);
    return factor_impl<float>(static_cast<float*>(A), lda, n, nb, ipiv, info, use_mfma, block_pivot, update_fp16, ws,
                              redone, host_flag, s);
}

static size_t perm_bytes(int n) { return ((sizeof(int) * (size_t)n + 255) / 256) * 256; }

size_t lu_solve_aux_bytes(int n, int elem_bytes)
{
    const size_t blocks = ((size_t)n + kSB - 1) / kSB;
    return perm_bytes(n) + blocks * 2 * kSB * kSB * (size_t)elem_bytes;
}

hipError_t lu_prepare_solve(const void* LU, int elem_bytes, int64_t lda, int n, const int* ipiv, void* aux,
                            hipStream_t s)
{
    // the swaps are sequential: composed on the host (n ints each way)
    std::vector<int> piv(n), perm(n);
    hipError_t e = hipMemcpyAsync(piv.data(), ipiv, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        const int p = piv[k];
        if (p < 0 || p >= n) return hipErrorInvalidValue;
        if (p != k) std::swap(perm[k], perm[p]);
    }
    e = hipMemcpyAsync(aux, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // perm is a local buffer
    if (e != hipSuccess) return e;
    const dim3 grid((n + kSB - 1) / kSB, 2);
    void* inv = static_cast<char*>(aux) + perm_bytes(n);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(block_inverse_kernel<double>, grid, dim3(kSB), 0, s, static_cast<const double*>(LU), lda,
                           n, static_cast<double*>(inv));
    else
        hipLaunchKernelGGL(block_inverse_kernel<float>, grid, dim3(kSB), 0, s, static_cast<const float*>(LU), lda, n,
                           static_cast<float*>(inv));
    return hipGetLastError();
}

template <typename T>
hipError_t solve_blocks_impl(const T* LU, int64_t lda, int n, const void* aux, double* x, double* tmp, hipStream_t s)
{
    const int* perm = static_cast<const int*>(aux);
    const T* inv = reinterpret_cast<const T*>(static_cast<const char*>(aux) + perm_bytes(n));
    const int nb = (n + kSB - 1) / kSB;
    hipLaunchKernelGGL(permute_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, perm, n, tmp);
    for (int b = 0; b < nb; ++b) {  // L z = P x: tmp -> x
        const int k0 = b * kSB, bs = std::min(kSB, n - k0);
        const int rows = n - k0 - bs;
        hipLaunchKernelGGL((sweep_block_kernel<T, 1>), dim3(std::max(1, (rows + 63) / 64)), dim3(256), 0, s, LU,
                           lda, n, inv, tmp, x, k0, bs);
    }
    for (int b = nb - 1; b >= 0; --b) {  // U y = z: x -> tmp
        const int k0 = b * kSB, bs = std::min(kSB, n - k0);
        hipLaunchKernelGGL((sweep_block_kernel<T, 0>), dim3(std::max(1, (k0 + 63) / 64)), dim3(256), 0, s, LU,
                           lda, n, inv, x, tmp, k0, bs);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_copy_vec(tmp, x, n, s);
}

hipError_t lu_solve_blocks(const void* LU, int elem_bytes, int64_t lda, int n, const void* aux, double* x,
                           double* tmp, hipStream_t s)
{
    if (elem_bytes == 8)
        return solve_blocks_impl<double>(static_cast<const double*>(LU), lda, n, aux, x, tmp, s);
    return solve_blocks_impl<float>(static_cast<const float*>(LU), lda, n, aux, x, tmp, s);
}

}  // namespace rf
