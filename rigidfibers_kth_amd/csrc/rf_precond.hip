// rf_precond.hip -- fiber neighbourhoods on gfx950: k nearest fibers by
// segment-segment distance, and the neighbour-coupled block-Jacobi
// (restricted additive Schwarz) preconditioner built on them.
//
// The diagonal blocks of the system matrix are exactly the identity
// (fortran/matrix_assembly.f90:308-311), so a per-fiber block-Jacobi
// preconditioner is the identity.  The ill-conditioning comes from fibers in
// near contact, whose mutual blocks are large.  For every fiber i the cluster
// C_i = {i} + its K nearest fibers is solved exactly:
//     z_i = [ (A_{C_i C_i})^{-1} r_{C_i} ]_i
// The 3M rows of the cluster inverse that belong to fiber i are precomputed
// once per assembly (Gauss-Jordan in LDS, one workgroup per fiber), so one
// application is a gather + small dense product per fiber.
#include "rf_internal.h"
#include "rf_precond.h"

#include <algorithm>

namespace rf {
namespace {

// Distance between the segments c1 + s t1 and c2 + u t2, |s|,|u| <= 1
// (closest points, clamped to both segments).
__device__ __forceinline__ double segment_distance(const double* c1, const double* t1, const double* c2,
                                                   const double* t2)
{
    const double w0[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
    const double a = t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2];
    const double b = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
    const double c = t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2];
    const double d = t1[0] * w0[0] + t1[1] * w0[1] + t1[2] * w0[2];
    const double e = t2[0] * w0[0] + t2[1] * w0[1] + t2[2] * w0[2];
    const double den = a * c - b * b;
    double s = (den > 1e-12) ? (b * e - c * d) / den : 0.0;
    s = fmin(1.0, fmax(-1.0, s));
    double u = (b * s + e) / c;
    u = fmin(1.0, fmax(-1.0, u));
    s = fmin(1.0, fmax(-1.0, (b * u - d) / a));
    const double dx = w0[0] + s * t1[0] - u * t2[0];
    const double dy = w0[1] + s * t1[1] - u * t2[1];
    const double dz = w0[2] + s * t1[2] - u * t2[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// One wavefront per fiber i: every lane keeps a sorted top-K list of its
// share of the sources, then K rounds of a wavefront min-reduction merge the
// lists (ties broken by the smaller index, so the result is deterministic).
template <int KMAX>
__global__ void __launch_bounds__(256) knn_kernel(int N, int K, const double* __restrict__ pos,
                                                  const double* __restrict__ orient, int* __restrict__ nbr,
                                                  double* __restrict__ nbr_dist)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= N) return;
    const double ci[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double ti[3] = {orient[3 * i], orient[3 * i + 1], orient[3 * i + 2]};
    double dl[KMAX];
    int il[KMAX];
#pragma unroll
    for (int q = 0; q < KMAX; ++q) {
        dl[q] = 1e300;
        il[q] = N;
    }
    for (int j = lane; j < N; j += 64) {
        if (j == i) continue;
        const double cj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
        const double tj[3] = {orient[3 * j], orient[3 * j + 1], orient[3 * j + 2]};
        double dv = segment_distance(ci, ti, cj, tj);
        int iv = j;
        if (dv < dl[KMAX - 1]) {
            // insertion into the sorted local list
#pragma unroll
            for (int q = 0; q < KMAX; ++q) {
                if (dv < dl[q]) {
                    const double td = dl[q];
                    const int ti2 = il[q];
                    dl[q] = dv;
                    il[q] = iv;
                    dv = td;
                    iv = ti2;
                }
            }
        }
    }
    int head = 0;
    for (int r = 0; r < K; ++r) {
        double hv = 1e300;
        int hi = N;
#pragma unroll
        for (int q = 0; q < KMAX; ++q)
            if (q == head) {
                hv = dl[q];
                hi = il[q];
            }
        double bv = hv;
        int bi = hi;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov < bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (hi == bi && hv == bv && head < KMAX) ++head;  // the winning lane pops its head
        if (lane == 0) {
            nbr[(size_t)i * K + r] = (bi < N) ? bi : i;
            nbr_dist[(size_t)i * K + r] = bv;
        }
    }
}

// One workgroup per fiber: Gauss-Jordan with partial pivoting on the
// transposed cluster matrix augmented with the first 3M unit vectors, which
// leaves in the augmented columns the rows of A_CC^{-1} that belong to fiber
// i.  B[i] is stored as [dim][bs] (bs = 3M, dim = bs (K+1)).
template <typename T>
__global__ void __launch_bounds__(256) ras_setup_kernel(int N, int K, int ldn, int bs, const int* __restrict__ nbr,
                                                        const T* __restrict__ A, int64_t lda,
                                                        double* __restrict__ B, int* __restrict__ info)
{
    extern __shared__ double sm[];
    const int dim = bs * (K + 1);
    const int ldm = dim + bs + 1;  // row stride: dim columns + bs augmented + pad
    double* Mt = sm;               // [dim][ldm]
    __shared__ int fib[16];
    __shared__ int piv;
    __shared__ double pivval;
    const int i = blockIdx.x;
    if (threadIdx.x == 0) {
        fib[0] = i;
        for (int c = 0; c < K; ++c) fib[c + 1] = nbr[(size_t)i * ldn + c];
    }
    __syncthreads();
    // Mt[r][c] = A_CC[c][r]; cluster row/col q -> global (fib[q/bs]*bs + q%bs)
    for (int e = threadIdx.x; e < dim * dim; e += blockDim.x) {
        const int rr = e % dim, cc = e / dim;  // A_CC row rr, column cc
        const int64_t grow = (int64_t)fib[rr / bs] * bs + rr % bs;
        const int64_t gcol = (int64_t)fib[cc / bs] * bs + cc % bs;
        Mt[cc * ldm + rr] = (double)A[gcol * lda + grow];
    }
    for (int e = threadIdx.x; e < dim * bs; e += blockDim.x) {
        const int rr = e / bs, q = e % bs;
        Mt[rr * ldm + dim + q] = (rr == q) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int W = dim + bs;
    for (int k = 0; k < dim; ++k) {
        if (threadIdx.x < 64) {
            double best = -1.0;
            int bi = dim;
            for (int r = k + threadIdx.x; r < dim; r += 64) {
                const double v = fabs(Mt[r * ldm + k]);
                if (v > best) {
                    best = v;
                    bi = r;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ov > best || (ov == best && oi < bi)) {
                    best = ov;
                    bi = oi;
                }
            }
            if (threadIdx.x == 0) piv = bi;
        }
        __syncthreads();
        const int p = piv;
        if (p != k && p < dim) {
            for (int c = threadIdx.x; c < W; c += blockDim.x) {
                const double t = Mt[k * ldm + c];
                Mt[k * ldm + c] = Mt[p * ldm + c];
                Mt[p * ldm + c] = t;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) pivval = Mt[k * ldm + k];
        __syncthreads();
        const double pv = pivval;
        if (pv == 0.0) {
            if (threadIdx.x == 0) atomicExch(info, i + 1);
            return;
        }
        const double rp = 1.0 / pv;
        // eliminate column k from every other row; row k is scaled last
        for (int e = threadIdx.x; e < dim * W; e += blockDim.x) {
            const int r = e / W, c = e % W;
            if (r == k || c == k) continue;
            const double f = Mt[r * ldm + k] * rp;
            Mt[r * ldm + c] -= f * Mt[k * ldm + c];
        }
        __syncthreads();
        for (int r = threadIdx.x; r < dim; r += blockDim.x)
            if (r != k) Mt[r * ldm + k] = 0.0;
        for (int c = threadIdx.x; c < W; c += blockDim.x) Mt[k * ldm + c] *= rp;
        __syncthreads();
    }
    // augmented column q = row q of A_CC^{-1}: B[i][c][q] = Mt[c][dim + q]
    double* Bi = B + (size_t)i * dim * bs;
    for (int e = threadIdx.x; e < dim * bs; e += blockDim.x) {
        const int c = e / bs, q = e % bs;
        Bi[e] = Mt[c * ldm + dim + q];
    }
}

// z_i = B_i r_{C_i}: thread per (fiber, row q)
__global__ void ras_apply_kernel(int N, int K, int ldn, int bs, const int* __restrict__ nbr, const double* __restrict__ B,
                                 const double* __restrict__ r, double* __restrict__ z)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * bs) return;
    const int i = t / bs, q = t - (t / bs) * bs;
    const int dim = bs * (K + 1);
    const double* Bi = B + (size_t)i * dim * bs + q;
    double s = 0.0;
    for (int c = 0; c <= K; ++c) {
        const int f = (c == 0) ? i : nbr[(size_t)i * ldn + c - 1];
        const double* rf = r + (size_t)f * bs;
        const double* bc = Bi + (size_t)c * bs * bs;
        for (int e = 0; e < bs; ++e) s += bc[(size_t)e * bs] * rf[e];
    }
    z[t] = s;
}

}  // namespace

hipError_t launch_knn(int N, int K, const double* pos, const double* orient, int* nbr, double* nbr_dist,
                      hipStream_t s)
{
    if (K < 1 || K > kMaxNeighbours || K > N - 1) return hipErrorInvalidValue;
    const int fibers_per_block = 4;
    hipLaunchKernelGGL(knn_kernel<kMaxNeighbours>, dim3((N + fibers_per_block - 1) / fibers_per_block),
                       dim3(64 * fibers_per_block), 0, s, N, K, pos, orient, nbr, nbr_dist);
    return hipGetLastError();
}

size_t ras_setup_lds(int K, int bs)
{
    const int dim = bs * (K + 1);
    return sizeof(double) * (size_t)dim * (dim + bs + 1);
}

hipError_t launch_ras_setup(int N, int K, int ldn, int bs, const int* nbr, const void* A, int elem_bytes,
                            int64_t lda, double* B, int* info, hipStream_t s)
{
    const size_t lds = ras_setup_lds(K, bs);
    if (K < 1 || K > kMaxRasNeighbours || K > ldn || lds > 160 * 1024) return hipErrorInvalidValue;
    if (elem_bytes == 8)
        hipLaunchKernelGGL(ras_setup_kernel<double>, dim3(N), dim3(256), lds, s, N, K, ldn, bs, nbr,
                           static_cast<const double*>(A), lda, B, info);
    else
        hipLaunchKernelGGL(ras_setup_kernel<float>, dim3(N), dim3(256), lds, s, N, K, ldn, bs, nbr,
                           static_cast<const float*>(A), lda, B, info);
    return hipGetLastError();
}

hipError_t launch_ras_apply(int N, int K, int ldn, int bs, const int* nbr, const double* B, const double* r,
                            double* z, hipStream_t s)
{
    const int tot = N * bs;
    hipLaunchKernelGGL(ras_apply_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, N, K, ldn, bs, nbr, B, r, z);
    return hipGetLastError();
}

}  // namespace rf
