// rf_linalg.hip -- HBM-bound vector kernels of the solver on gfx950:
// column-major GEMV (the GMRES matvec of fortran/coeffs_and_vels.f90:299,483
// and the velocity-operator product), deterministic multi-dot reductions and
// Gram-Schmidt updates of the Arnoldi basis.
//
// All arithmetic is FP64.  The matrix may be stored as double or float; the
// float case (mixed precision) halves the bytes streamed from HBM while the
// products and sums stay FP64.
#include "rf_internal.h"

#include <algorithm>

namespace rf {
namespace {

constexpr int kGemvThreads = 256;  // 4 wavefronts per workgroup
constexpr int kGemvUnroll = 8;     // columns in flight per wavefront

// 16-byte vectors of the matrix, read with non-temporal (streaming) loads:
// the matrix is read once per GEMV and should not evict the vectors.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
    using type = d2v;
    static constexpr int n = 2;
};
template <>
struct Vec16<float> {
    using type = f4v;
    static constexpr int n = 4;
};

template <typename V, int N>
__device__ __forceinline__ void unpack(const V& v, double (&o)[N])
{
#pragma unroll
    for (int r = 0; r < N; ++r) o[r] = (double)v[r];
}

// parts[split][row] = sum_{c in split} A[row + c*ld] * x[c]
// Workgroup: 64*RPL consecutive rows x one column split.  Wavefront w of the
// workgroup takes columns c0+w, c0+w+4, ...; each lane loads 16 bytes (RPL
// rows) per column, so one wavefront load is a 1 KiB contiguous segment.
// With xs.nrm2 the vector is x / sqrt(*xs.nrm2), formed exactly as
// scale_copy_kernel forms it, and the workgroups of the first row block
// store it to xs.out (each column once): the next Arnoldi vector and its
// product in one pass.
struct XScale {
    const double* nrm2;
    double* out;
};

template <typename T>
__global__ void __launch_bounds__(kGemvThreads) gemv_parts_kernel(int cols, int64_t ld, const T* __restrict__ A,
                                                                  const double* __restrict__ x,
                                                                  double* __restrict__ parts, int64_t pstride,
                                                                  int cols_per_split, XScale xs)
{
    using V = typename Vec16<T>::type;
    constexpr int RPL = Vec16<T>::n;
    constexpr int ROWS = 64 * RPL;
    __shared__ double red[4][ROWS];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t row0 = (int64_t)blockIdx.x * ROWS + lane * RPL;
    const int c0 = blockIdx.y * cols_per_split;
    const int c1 = min(cols, c0 + cols_per_split);
    double acc[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) acc[r] = 0.0;
    const bool scaled = xs.nrm2 != nullptr;
    const double inv = scaled ? 1.0 / sqrt(*xs.nrm2) : 1.0;
    const bool store_x = scaled && xs.out != nullptr && blockIdx.x == 0 && lane == 0;
    const T* Ab = A + row0;
    int c = c0 + wave;
    for (; c + 4 * (kGemvUnroll - 1) < c1; c += 4 * kGemvUnroll) {
        V v[kGemvUnroll];
        double xv[kGemvUnroll];
#pragma unroll
        for (int u = 0; u < kGemvUnroll; ++u) {
            v[u] = __builtin_nontemporal_load(reinterpret_cast<const V*>(Ab + (int64_t)(c + 4 * u) * ld));
            xv[u] = scaled ? x[c + 4 * u] * inv : x[c + 4 * u];
        }
        if (store_x) {
#pragma unroll
            for (int u = 0; u < kGemvUnroll; ++u) xs.out[c + 4 * u] = xv[u];
        }
#pragma unroll
        for (int u = 0; u < kGemvUnroll; ++u) {
            double e[RPL];
            unpack(v[u], e);
#pragma unroll
            for (int r = 0; r < RPL; ++r) acc[r] += e[r] * xv[u];
        }
    }
    for (; c < c1; c += 4) {
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(Ab + (int64_t)c * ld));
        const double xv = scaled ? x[c] * inv : x[c];
        if (store_x) xs.out[c] = xv;
        double e[RPL];
        unpack(v, e);
#pragma unroll
        for (int r = 0; r < RPL; ++r) acc[r] += e[r] * xv;
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) red[wave][lane * RPL + r] = acc[r];
    __syncthreads();
    for (int rr = threadIdx.x; rr < ROWS; rr += kGemvThreads) {
        const double s = (red[0][rr] + red[1][rr]) + (red[2][rr] + red[3][rr]);
        parts[(int64_t)blockIdx.y * pstride + (int64_t)blockIdx.x * ROWS + rr] = s;
    }
}

// y[i] = sum_s parts[s][i]  (mode 0)  or  b[i] - sum_s parts[s][i]  (mode 1).
// Eight lanes per row: lane q sums the splits q, q+8, ... (all its loads in
// flight together), and the eight partial sums meet in a fixed xor tree, so
// the result is deterministic.
constexpr int kSplitLanes = 8;

__global__ void reduce_splits_kernel(const double* __restrict__ parts, int splits, int64_t stride, int n,
                                     const double* __restrict__ b, int mode, double* __restrict__ y)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(t / kSplitLanes), q = (int)(t % kSplitLanes);
    double s = 0.0;
    if (i < n) {
        int k = q;
        for (; k + 3 * kSplitLanes < splits; k += 4 * kSplitLanes) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = parts[(int64_t)(k + u * kSplitLanes) * stride + i];
#pragma unroll
            for (int u = 0; u < 4; ++u) s += v[u];
        }
        for (; k < splits; k += kSplitLanes) s += parts[(int64_t)k * stride + i];
    }
#pragma unroll
    for (int off = 1; off < kSplitLanes; off <<= 1) s += __shfl_xor(s, off, 64);
    if (q == 0 && i < n) y[i] = (mode == 1) ? b[i] - s : s;
}

constexpr int kDotThreads = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
    }
    __syncthreads();
    return s;  // valid in thread 0
}

// Block (bx, j): partial of dot(V[:,j], w) (j < k) or dot(w,w) (j == k) over a
// row chunk.  The last workgroup to finish sums the partials in a fixed order
// (deterministic) into out[0..k] and resets the counter.
__global__ void __launch_bounds__(kDotThreads) multidot_kernel(const double* __restrict__ V, int64_t ldv, int k,
                                                               const double* __restrict__ w, int n,
                                                               double* __restrict__ parts,
                                                               unsigned int* counter, double* __restrict__ out)
{
    __shared__ double sh[kDotThreads / 64];
    __shared__ bool last;
    const int j = blockIdx.y;
    const int nbx = gridDim.x;
    const double* a = (j < k) ? V + (int64_t)j * ldv : w;
    double s = 0.0;
    for (int i = blockIdx.x * kDotThreads + threadIdx.x; i < n; i += nbx * kDotThreads) s += a[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        parts[(int64_t)j * nbx + blockIdx.x] = s;
        __threadfence();
        const unsigned int total = gridDim.x * gridDim.y;
        const unsigned int prev = atomicAdd(counter, 1u);
        last = (prev == total - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        for (int jj = threadIdx.x; jj <= k; jj += kDotThreads) {
            double t = 0.0;
            for (int b = 0; b < nbx; ++b)
                t += __hip_atomic_load(&parts[(int64_t)jj * nbx + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            out[jj] = t;
        }
        if (threadIdx.x == 0) atomicExch(counter, 0u);
    }
}

// w -= V[:, :k] h
__global__ void orth_update_kernel(const double* __restrict__ V, int64_t ldv, int k, const double* __restrict__ h,
                                   double* __restrict__ w, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int j = 0; j < k; ++j) s -= V[(int64_t)j * ldv + i] * h[j];
    w[i] = s;
}

// Fused w -= V[:, :k] h and the dots that follow it: out[j] = V[:, j]^T w for
// j < nd (nd = k or 0) and out[nd] = w^T w, with w updated in the same pass.
// Rows, partial sums and the last workgroup's fixed-order reduction follow
// orth_update_kernel then multidot_kernel exactly (same grid-stride rows per
// thread, same wave and block sums, same partial layout), so the results are
// bit-identical to the two-kernel sequence with one launch less.  k <= MAXK.
// With hc.dst the last workgroup then does what to_host_kernel does: copies
// hc.words words from hc.src (out included) to host memory and stores the
// sequence word after them at system scope.
template <int MAXK>
__global__ void __launch_bounds__(kDotThreads) orth_dots_kernel(const double* __restrict__ V, int64_t ldv, int k,
                                                                const double* __restrict__ h,
                                                                double* __restrict__ w, int n, int nd,
                                                                double* __restrict__ parts,
                                                                unsigned int* counter, double* __restrict__ out,
                                                                HostCopy hc)
{
    __shared__ double hs[MAXK];
    __shared__ double sh[MAXK + 1][kDotThreads / 64];
    __shared__ bool last;
    if (threadIdx.x < MAXK) hs[threadIdx.x] = (threadIdx.x < k) ? h[threadIdx.x] : 0.0;
    __syncthreads();
    double acc[MAXK + 1];
#pragma unroll
    for (int j = 0; j <= MAXK; ++j) acc[j] = 0.0;
    const int nbx = gridDim.x;
    for (int i = blockIdx.x * kDotThreads + threadIdx.x; i < n; i += nbx * kDotThreads) {
        double v[MAXK];
#pragma unroll
        for (int j = 0; j < MAXK; ++j) v[j] = (j < k) ? V[(int64_t)j * ldv + i] : 0.0;
        double x = w[i];
#pragma unroll
        for (int j = 0; j < MAXK; ++j)
            if (j < k) x -= v[j] * hs[j];
        w[i] = x;
#pragma unroll
        for (int j = 0; j < MAXK; ++j)
            if (j < nd) acc[j] += v[j] * x;
        acc[MAXK] += x * x;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j <= MAXK; ++j) {
        if (j < nd || j == MAXK) {
            const double t = wave_sum(acc[j]);
            if (lane == 0) sh[j][wv] = t;
        }
    }
    __syncthreads();
    if (threadIdx.x <= nd) {  // slot nd carries the norm
        const int src = (threadIdx.x == nd) ? MAXK : threadIdx.x;
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sh[src][q];
        parts[(int64_t)threadIdx.x * nbx + blockIdx.x] = t;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int prev = atomicAdd(counter, 1u);
        last = (prev == (unsigned int)nbx - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        for (int jj = threadIdx.x; jj <= nd; jj += kDotThreads) {
            double t = 0.0;
            for (int b = 0; b < nbx; ++b)
                t += __hip_atomic_load(&parts[(int64_t)jj * nbx + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            out[jj] = t;
        }
        if (threadIdx.x == 0) atomicExch(counter, 0u);
        if (hc.dst) {
            __threadfence();
            __syncthreads();
            for (int i = threadIdx.x; i < hc.words; i += kDotThreads)
                hc.dst[i] = __hip_atomic_load(&hc.src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence_system();
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(hc.seq_word, hc.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// v = w / sqrt(nrm2[idx])
__global__ void scale_copy_kernel(const double* __restrict__ w, const double* __restrict__ nrm2, int idx, int n,
                                  double* __restrict__ v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double inv = 1.0 / sqrt(nrm2[idx]);
    v[i] = w[i] * inv;
}

// xc = x + V[:, :k] y with y passed by value (the host's least-squares
// solution: a host-to-device copy before the launch would stall the queue)
constexpr int kYArg = 64;
struct YArg {
    double v[kYArg];
};

__global__ void vy_plus_x_kernel(const double* __restrict__ V, int64_t ldv, int k, YArg y, const double* x, int n,
                                 double* xc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += V[(int64_t)j * ldv + i] * y.v[j];
    xc[i] = x ? x[i] + s : s;
}

__global__ void copy_vec_kernel(const double* __restrict__ src, double* __restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void f64_to_f32_kernel(const double* __restrict__ src, float* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

// A few words into coherent pinned host memory, by one workgroup: the
// runtime's device-to-host copy ends with a system-scope release that stalls
// the queue ~15 us, once per GMRES iteration.
// With seq_word, the copy ends with `seq` stored there at system scope after
// every thread's data stores have been fenced at system scope, so a host that
// reads the expected sequence number reads the data of this copy.
__global__ void to_host_kernel(const unsigned* __restrict__ src, unsigned* dst, int words, unsigned* seq_word,
                               unsigned seq)
{
    for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
    if (seq_word) {
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(seq_word, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace

hipError_t launch_to_host(const void* src, void* host_dst, size_t bytes, hipStream_t s, unsigned* host_seq_word,
                          unsigned seq)
{
    if (bytes % sizeof(unsigned) != 0 || bytes > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(to_host_kernel, dim3(1), dim3(256), 0, s, static_cast<const unsigned*>(src),
                       static_cast<unsigned*>(host_dst), (int)(bytes / sizeof(unsigned)), host_seq_word, seq);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double* src, float* dst, size_t n, hipStream_t s)
{
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s)
{
    const size_t n4 = bytes / sizeof(double4);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, static_cast<const double4*>(src),
                       static_cast<double4*>(dst), n4);
    return hipGetLastError();
}

GemvPlan plan_gemv(int rows, int cols, int64_t ld, int elem_bytes, int num_cus)
{
    GemvPlan p;
    p.rows = rows;
    p.cols = cols;
    p.ld = ld;
    p.elem_bytes = elem_bytes;
    const int rows_per_block = 64 * (16 / elem_bytes);
    p.row_blocks = (int)((ld + rows_per_block - 1) / rows_per_block);
    // about 24 workgroups per CU (measured best for n = 15000..60000 with
    // streaming loads), but keep >= 32 columns per wavefront
    const int target = std::max(1, num_cus) * 24;
    int splits = (target + p.row_blocks - 1) / p.row_blocks;
    splits = std::min(splits, std::max(1, cols / 128));
    p.splits = std::max(1, std::min(splits, 256));
    return p;
}

size_t gemv_parts_elems(const GemvPlan& p)
{
    const int rows_per_block = 64 * (16 / p.elem_bytes);
    return (size_t)p.splits * (size_t)p.row_blocks * rows_per_block;
}

hipError_t launch_gemv(const GemvPlan& p, const void* A, const double* x, double* parts, double* y, const double* b,
                       int mode, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, const double* x_nrm2, double* x_out)
{
    const XScale xs{x_nrm2, x_out};
    const int rows_per_block = 64 * (16 / p.elem_bytes);
    if (p.ld % rows_per_block != 0) return hipErrorInvalidValue;  // loads assume whole row blocks
    const int64_t pstride = (int64_t)p.row_blocks * rows_per_block;
    const int cps = (p.cols + p.splits - 1) / p.splits;
    dim3 grid(p.row_blocks, p.splits);
    if (ev0) {
        hipError_t e0 = hipEventRecord(ev0, s);
        if (e0 != hipSuccess) return e0;
    }
    if (p.elem_bytes == 8)
        hipLaunchKernelGGL(gemv_parts_kernel<double>, grid, dim3(kGemvThreads), 0, s, p.cols, p.ld,
                           static_cast<const double*>(A), x, parts, pstride, cps, xs);
    else
        hipLaunchKernelGGL(gemv_parts_kernel<float>, grid, dim3(kGemvThreads), 0, s, p.cols, p.ld,
                           static_cast<const float*>(A), x, parts, pstride, cps, xs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev1) {
        e = hipEventRecord(ev1, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(reduce_splits_kernel, dim3((p.rows * kSplitLanes + 255) / 256), dim3(256), 0, s, parts,
                       p.splits, pstride, p.rows, b, mode, y);
    return hipGetLastError();
}

hipError_t launch_reduce_splits(const double* parts, int splits, int64_t stride, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_splits_kernel, dim3((n * kSplitLanes + 255) / 256), dim3(256), 0, s, parts, splits,
                       stride, n, nullptr, 0, out);
    return hipGetLastError();
}

int multidot_blocks(int n)
{
    return std::max(1, std::min(kMaxDotBlocks, (n + 4 * kDotThreads - 1) / (4 * kDotThreads)));
}

hipError_t launch_multidot(const double* V, int64_t ldv, int k, const double* w, int n, double* parts,
                           unsigned int* counter, double* out, hipStream_t s)
{
    dim3 grid(multidot_blocks(n), k + 1);
    hipLaunchKernelGGL(multidot_kernel, grid, dim3(kDotThreads), 0, s, V, ldv, k, w, n, parts, counter, out);
    return hipGetLastError();
}

hipError_t launch_orth_dots(const double* V, int64_t ldv, int k, const double* h, double* w, int n, int with_dots,
                            double* parts, unsigned int* counter, double* out, hipStream_t s, const HostCopy& hc)
{
    const int nd = with_dots ? k : 0;
    const dim3 grid(multidot_blocks(n)), block(kDotThreads);
    if (hc.dst && (hc.words < 0 || hc.words > 16384 || !hc.src || !hc.seq_word)) return hipErrorInvalidValue;
    if (k <= 16) {
        hipLaunchKernelGGL(orth_dots_kernel<16>, grid, block, 0, s, V, ldv, k, h, w, n, nd, parts, counter, out, hc);
        return hipGetLastError();
    }
    if (k <= 32) {
        hipLaunchKernelGGL(orth_dots_kernel<32>, grid, block, 0, s, V, ldv, k, h, w, n, nd, parts, counter, out, hc);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(orth_update_kernel, dim3((n + 255) / 256), dim3(256), 0, s, V, ldv, k, h, w, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_multidot(with_dots ? V : nullptr, ldv, nd, w, n, parts, counter, out, s);
    if (e != hipSuccess || !hc.dst) return e;
    return launch_to_host(hc.src, hc.dst, sizeof(unsigned) * (size_t)hc.words, s, hc.seq_word, hc.seq);
}

hipError_t launch_scale_copy(const double* w, const double* nrm2, int idx, int n, double* v, hipStream_t s)
{
    hipLaunchKernelGGL(scale_copy_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w, nrm2, idx, n, v);
    return hipGetLastError();
}

hipError_t launch_vy_plus_x(const double* V, int64_t ldv, int k, const double* y_host, const double* x, int n,
                            double* xc, hipStream_t s)
{
    for (int c = 0; c < k || c == 0; c += kYArg) {  // chunks of 64 basis vectors
        YArg y{};
        const int kc = std::min(kYArg, k - c);
        for (int j = 0; j < kc; ++j) y.v[j] = y_host[c + j];
        hipLaunchKernelGGL(vy_plus_x_kernel, dim3((n + 255) / 256), dim3(256), 0, s, V + (int64_t)c * ldv, ldv,
                           std::max(0, kc), y, c == 0 ? x : xc, n, xc);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (kc <= 0) break;
    }
    return hipSuccess;
}

hipError_t launch_copy_vec(const double* src, double* dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(copy_vec_kernel, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

}  // namespace rf
