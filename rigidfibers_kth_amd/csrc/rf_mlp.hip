// rf_mlp.hip -- the learned initial guess on gfx950: per-fiber features from
// the 16 nearest fibers, a 113-128-128-128-15 tanh MLP (FP32) that predicts
// the fiber's force coefficients in its own frame, and its training
// (forward, backward, mean-squared-error loss, Adam), all hand-written.
//
// Local frame of fiber i: e1 = t_i, e2 = normalize(z x t_i) (x x t_i when t_i
// is parallel to z), e3 = e1 x e2.  Features (113): for each of the 16
// nearest fibers j (segment-segment distance, ascending) the relative centre
// (x_j - x_i)/2, the orientation t_j with the sign that makes t_j.e1 >= 0
// (a fiber is the same segment for +t and -t), both in the local frame, and
// the proximity 0.1/(d_ij + 0.1); then t_i.z.  Targets (15): the coefficients
// a_i^k (k = 1..M, M = 5) in the local frame.
//
// Weight blob (floats): in_mean[113] in_std[113] W1[128x113] b1[128]
// W2[128x128] b2[128] W3[128x128] b3[128] W4[15x128] b4[15] out_mean[15]
// out_std[15].  W is row-major [out][in].
#include "rf_internal.h"
#include "rf_mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace rf {
namespace {

// ---- features / frames ---------------------------------------------------

__device__ __forceinline__ void local_frame(const double* t, double* e1, double* e2, double* e3)
{
    e1[0] = t[0];
    e1[1] = t[1];
    e1[2] = t[2];
    // e2 = z x t  (or x x t)
    double a0 = -t[1], a1 = t[0], a2 = 0.0;
    double nrm = sqrt(a0 * a0 + a1 * a1);
    if (nrm < 1e-6) {
        a0 = 0.0;
        a1 = -t[2];
        a2 = t[1];
        nrm = sqrt(a1 * a1 + a2 * a2);
    }
    e2[0] = a0 / nrm;
    e2[1] = a1 / nrm;
    e2[2] = a2 / nrm;
    e3[0] = e1[1] * e2[2] - e1[2] * e2[1];
    e3[1] = e1[2] * e2[0] - e1[0] * e2[2];
    e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
}

__global__ void features_kernel(int N, int ldn, const double* __restrict__ pos, const double* __restrict__ orient,
                                const int* __restrict__ nbr, const double* __restrict__ nbr_dist,
                                float* __restrict__ X)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double e1[3], e2[3], e3[3];
    local_frame(orient + 3 * i, e1, e2, e3);
    float* xi = X + (size_t)i * kMlpIn;
    const double* ci = pos + 3 * i;
    for (int r = 0; r < kMlpNeighbours; ++r) {
        const int j = nbr[(size_t)i * ldn + r];
        const double d = nbr_dist[(size_t)i * ldn + r];
        const double dx = pos[3 * j] - ci[0], dy = pos[3 * j + 1] - ci[1], dz = pos[3 * j + 2] - ci[2];
        const double* tj = orient + 3 * j;
        double u1 = tj[0] * e1[0] + tj[1] * e1[1] + tj[2] * e1[2];
        double u2 = tj[0] * e2[0] + tj[1] * e2[1] + tj[2] * e2[2];
        double u3 = tj[0] * e3[0] + tj[1] * e3[1] + tj[2] * e3[2];
        const double sg = (u1 > 0.0 || (u1 == 0.0 && (u2 > 0.0 || (u2 == 0.0 && u3 >= 0.0)))) ? 1.0 : -1.0;
        float* f = xi + 7 * r;
        f[0] = (float)(0.5 * (dx * e1[0] + dy * e1[1] + dz * e1[2]));
        f[1] = (float)(0.5 * (dx * e2[0] + dy * e2[1] + dz * e2[2]));
        f[2] = (float)(0.5 * (dx * e3[0] + dy * e3[1] + dz * e3[2]));
        f[3] = (float)(sg * u1);
        f[4] = (float)(sg * u2);
        f[5] = (float)(sg * u3);
        f[6] = (float)(0.1 / (d + 0.1));
    }
    xi[7 * kMlpNeighbours] = (float)e1[2];
}

// coefficients (global, fiber-major 3M) <-> local frame (M=5 -> 15 values)
__global__ void to_local_kernel(int N, int M, const double* __restrict__ orient, const double* __restrict__ x,
                                float* __restrict__ Y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double e1[3], e2[3], e3[3];
    local_frame(orient + 3 * i, e1, e2, e3);
    for (int k = 0; k < M; ++k) {
        const double* a = x + (size_t)i * 3 * M + 3 * k;
        Y[(size_t)i * 3 * M + 3 * k + 0] = (float)(a[0] * e1[0] + a[1] * e1[1] + a[2] * e1[2]);
        Y[(size_t)i * 3 * M + 3 * k + 1] = (float)(a[0] * e2[0] + a[1] * e2[1] + a[2] * e2[2]);
        Y[(size_t)i * 3 * M + 3 * k + 2] = (float)(a[0] * e3[0] + a[1] * e3[1] + a[2] * e3[2]);
    }
}

__global__ void to_global_kernel(int N, int M, const double* __restrict__ orient, const float* __restrict__ Y,
                                 const float* __restrict__ out_mean, const float* __restrict__ out_std,
                                 double* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double e1[3], e2[3], e3[3];
    local_frame(orient + 3 * i, e1, e2, e3);
    for (int k = 0; k < M; ++k) {
        const int q = 3 * k;
        const float* y = Y + (size_t)i * kMlpOut + q;
        const double l1 = (double)y[0] * out_std[q] + out_mean[q];
        const double l2 = (double)y[1] * out_std[q + 1] + out_mean[q + 1];
        const double l3 = (double)y[2] * out_std[q + 2] + out_mean[q + 2];
        double* a = x + (size_t)i * 3 * M + q;
        a[0] = l1 * e1[0] + l2 * e2[0] + l3 * e3[0];
        a[1] = l1 * e1[1] + l2 * e2[1] + l3 * e3[1];
        a[2] = l1 * e1[2] + l2 * e2[2] + l3 * e3[2];
    }
}

// ---- dense layers ----------------------------------------------------------
// C[m][n] (+)= sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak], B(k,n) =
// B[k*sbk + n*sbn]; 64x64 tiles per workgroup, K staged 16 at a time in LDS.
// epi: 0 plain store, 1 add bias b[n] then tanh, 2 add bias, 3 accumulate.
constexpr int kT = 64, kK = 16;

__global__ void __launch_bounds__(256) gemm_kernel(int Mr, int Nc, int Kd, const float* __restrict__ A, int64_t sam,
                                                   int64_t sak, const float* __restrict__ B, int64_t sbk, int64_t sbn,
                                                   float* __restrict__ C, int64_t ldc, const float* __restrict__ bias,
                                                   int epi)
{
    __shared__ float As[kK][kT + 1];
    __shared__ float Bs[kK][kT + 1];
    const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < Kd; k0 += kK) {
        for (int e = threadIdx.x; e < kK * kT; e += 256) {
            const int kk = e / kT, mm = e % kT;
            const int m = m0 + mm, k = k0 + kk;
            As[kk][mm] = (m < Mr && k < Kd) ? A[(int64_t)m * sam + (int64_t)k * sak] : 0.f;
            const int n = n0 + mm;
            Bs[kk][mm] = (n < Nc && k < Kd) ? B[(int64_t)k * sbk + (int64_t)n * sbn] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = As[kk][ty * 4 + u];
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
            const int m = m0 + ty * 4 + u, n = n0 + tx * 4 + v;
            if (m >= Mr || n >= Nc) continue;
            float val = acc[u][v];
            float* c = C + (int64_t)m * ldc + n;
            if (epi == 1) val = tanhf(val + bias[n]);
            else if (epi == 2) val = val + bias[n];
            else if (epi == 3) val = *c + val;
            *c = val;
        }
}

// (x - mean)/std per feature
__global__ void normalize_kernel(int B, int I, const float* __restrict__ X, const float* __restrict__ mean,
                                 const float* __restrict__ sd, float* __restrict__ Z)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * I) return;
    const int f = (int)(t % I);
    Z[t] = (X[t] - mean[f]) / sd[f];
}

// MSE loss on normalised targets: d = 2 (yhat - (y - mean)/sd) / (B*O);
// per-block partial losses.
__global__ void mse_kernel(int B, int O, const float* __restrict__ yhat, const float* __restrict__ Y,
                           const float* __restrict__ mean, const float* __restrict__ sd, float* __restrict__ d,
                           float* __restrict__ partial)
{
    __shared__ float sh[256];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (t < (int64_t)B * O) {
        const int o = (int)(t % O);
        const float diff = yhat[t] - (Y[t] - mean[o]) / sd[o];
        d[t] = 2.f * diff / (float)((int64_t)B * O);
        l = diff * diff;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// delta *= (1 - h^2)  (tanh derivative from the stored activation)
__global__ void tanh_back_kernel(int64_t n, const float* __restrict__ h, float* __restrict__ delta)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) delta[t] *= (1.f - h[t] * h[t]);
}

// db[o] = sum_b delta[b][o]
__global__ void bias_grad_kernel(int B, int O, const float* __restrict__ delta, float* __restrict__ db)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= O) return;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += delta[(int64_t)b * O + o];
    db[o] = s;
}

__global__ void adam_kernel(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

__global__ void gather_rows_kernel(int B, int W, const float* __restrict__ src, const int64_t* __restrict__ idx,
                                   float* __restrict__ dst)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * W) return;
    const int64_t b = t / W, c = t % W;
    dst[t] = src[idx[b] * W + c];
}

// acc[0] += sum of partial[0..n)  (one block)
__global__ void accumulate_kernel(int64_t n, const float* __restrict__ partial, double* __restrict__ acc)
{
    __shared__ double sh[256];
    double t = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) t += partial[i];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += sh[0];
}

inline dim3 grid1(int64_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

hipError_t gemm(int Mr, int Nc, int Kd, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk,
                int64_t sbn, float* C, int64_t ldc, const float* bias, int epi, hipStream_t s)
{
    dim3 grid((Nc + kT - 1) / kT, (Mr + kT - 1) / kT);
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, Mr, Nc, Kd, A, sam, sak, B, sbk, sbn, C, ldc, bias, epi);
    return hipGetLastError();
}

}  // namespace

MlpLayout mlp_layout()
{
    MlpLayout L;
    int64_t o = 0;
    L.in_mean = o;
    o += kMlpIn;
    L.in_std = o;
    o += kMlpIn;
    const int dims[5] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};
    L.first_param = o;
    for (int l = 0; l < 4; ++l) {
        L.W[l] = o;
        o += (int64_t)dims[l + 1] * dims[l];
        L.b[l] = o;
        o += dims[l + 1];
    }
    L.num_params = o - L.first_param;
    L.out_mean = o;
    o += kMlpOut;
    L.out_std = o;
    o += kMlpOut;
    L.total = o;
    return L;
}

hipError_t mlp_features(int N, int ldn, const double* pos, const double* orient, const int* nbr,
                        const double* nbr_dist, float* X, hipStream_t s)
{
    hipLaunchKernelGGL(features_kernel, grid1(N, 128), dim3(128), 0, s, N, ldn, pos, orient, nbr, nbr_dist, X);
    return hipGetLastError();
}

hipError_t mlp_targets(int N, int M, const double* orient, const double* x, float* Y, hipStream_t s)
{
    hipLaunchKernelGGL(to_local_kernel, grid1(N, 128), dim3(128), 0, s, N, M, orient, x, Y);
    return hipGetLastError();
}

// Forward pass of B rows of X; activations H[0..3] ([B][128] each) and the
// normalised output Yhat [B][15] are kept for the backward pass.
hipError_t mlp_forward(const float* w, int B, const float* X, MlpBuffers& buf, hipStream_t s)
{
    const MlpLayout L = mlp_layout();
    hipLaunchKernelGGL(normalize_kernel, grid1((int64_t)B * kMlpIn), dim3(256), 0, s, B, kMlpIn, X, w + L.in_mean,
                       w + L.in_std, buf.Z);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int dims[5] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};
    const float* in = buf.Z;
    for (int l = 0; l < 4; ++l) {
        float* out = (l < 3) ? buf.H[l] : buf.Yhat;
        e = gemm(B, dims[l + 1], dims[l], in, dims[l], 1, w + L.W[l], 1, dims[l], out, dims[l + 1], w + L.b[l],
                 (l < 3) ? 1 : 2, s);
        if (e != hipSuccess) return e;
        in = out;
    }
    return hipSuccess;
}

// Backward pass given dYhat in buf.D[3] ([B][15]); gradients into g (same
// offsets as the weight blob; normalisation entries untouched).
hipError_t mlp_backward(const float* w, int B, MlpBuffers& buf, float* g, hipStream_t s)
{
    const MlpLayout L = mlp_layout();
    const int dims[5] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};
    hipError_t e;
    for (int l = 3; l >= 0; --l) {
        const float* delta = buf.D[l];  // [B][dims[l+1]]
        const float* in = (l == 0) ? buf.Z : buf.H[l - 1];
        // dW[o][i] = sum_b delta[b][o] in[b][i]
        e = gemm(dims[l + 1], dims[l], B, delta, 1, dims[l + 1], in, dims[l], 1, g + L.W[l], dims[l], nullptr, 0, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(bias_grad_kernel, grid1(dims[l + 1], 128), dim3(128), 0, s, B, dims[l + 1], delta,
                           g + L.b[l]);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (l > 0) {
            // delta_prev[b][i] = sum_o delta[b][o] W[o][i], then * tanh'
            float* dprev = buf.D[l - 1];
            e = gemm(B, dims[l], dims[l + 1], delta, dims[l + 1], 1, w + L.W[l], dims[l], 1, dprev, dims[l], nullptr,
                     0, s);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(tanh_back_kernel, grid1((int64_t)B * dims[l]), dim3(256), 0, s, (int64_t)B * dims[l],
                               buf.H[l - 1], dprev);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

hipError_t mlp_loss(const float* w, int B, const float* Y, MlpBuffers& buf, hipStream_t s)
{
    const MlpLayout L = mlp_layout();
    const int64_t tot = (int64_t)B * kMlpOut;
    hipLaunchKernelGGL(mse_kernel, grid1(tot), dim3(256), 0, s, B, kMlpOut, buf.Yhat, Y, w + L.out_mean, w + L.out_std,
                       buf.D[3], buf.partial);
    return hipGetLastError();
}

int64_t mlp_loss_blocks(int B) { return ((int64_t)B * kMlpOut + 255) / 256; }

hipError_t mlp_to_global(int N, int M, const double* orient, const float* w, const float* Yhat, double* x,
                         hipStream_t s)
{
    const MlpLayout L = mlp_layout();
    hipLaunchKernelGGL(to_global_kernel, grid1(N, 128), dim3(128), 0, s, N, M, orient, Yhat, w + L.out_mean,
                       w + L.out_std, x);
    return hipGetLastError();
}

hipError_t mlp_adam(int n, float* w, const float* g, float* m, float* v, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, grid1(n), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t mlp_gather(int B, int W, const float* src, const int64_t* idx, float* dst, hipStream_t s)
{
    hipLaunchKernelGGL(gather_rows_kernel, grid1((int64_t)B * W), dim3(256), 0, s, B, W, src, idx, dst);
    return hipGetLastError();
}


namespace {
inline int64_t pad64(int64_t x) { return (x + 63) / 64 * 64; }
}  // namespace

int64_t mlp_buffer_floats(int rows)
{
    const int64_t r = rows;
    return pad64(r * kMlpIn) + 6 * pad64(r * kMlpHidden) + 2 * pad64(r * kMlpOut) + pad64(mlp_loss_blocks(rows));
}

void mlp_buffer_bind(float* base, int rows, MlpBuffers* b)
{
    const int64_t r = rows;
    float* q = base;
    b->rows = rows;
    b->Z = q;
    q += pad64(r * kMlpIn);
    for (int l = 0; l < 3; ++l) {
        b->H[l] = q;
        q += pad64(r * kMlpHidden);
    }
    b->Yhat = q;
    q += pad64(r * kMlpOut);
    for (int l = 0; l < 3; ++l) {
        b->D[l] = q;
        q += pad64(r * kMlpHidden);
    }
    b->D[3] = q;
    q += pad64(r * kMlpOut);
    b->partial = q;
}

hipError_t mlp_accumulate_loss(int B, const MlpBuffers& buf, double* acc, hipStream_t s)
{
    hipLaunchKernelGGL(accumulate_kernel, dim3(1), dim3(256), 0, s, mlp_loss_blocks(B), buf.partial, acc);
    return hipGetLastError();
}

// Glorot-uniform weights from SplitMix64 ((u >> 11) * 2^-53), zero biases.
void mlp_init_weights(uint64_t seed, float* w)
{
    const MlpLayout L = mlp_layout();
    uint64_t state = seed;
    auto next = [&state]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const int dims[5] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};
    for (int l = 0; l < 4; ++l) {
        const double lim = std::sqrt(6.0 / (dims[l] + dims[l + 1]));
        for (int64_t t = 0; t < (int64_t)dims[l + 1] * dims[l]; ++t) {
            const double u = (double)(next() >> 11) * 0x1.0p-53;
            w[L.W[l] + t] = (float)((2.0 * u - 1.0) * lim);
        }
        for (int t = 0; t < dims[l + 1]; ++t) w[L.b[l] + t] = 0.f;
    }
}

// Column means and deviations (floored) of the samples into the blob.
void mlp_set_standardisation(const float* X, const float* Y, int64_t count, float* w)
{
    const MlpLayout L = mlp_layout();
    auto stats = [count](const float* D, int cols, float* mean, float* sd, double floor_sd) {
        for (int c = 0; c < cols; ++c) {
            double s1 = 0.0, s2 = 0.0;
            for (int64_t r = 0; r < count; ++r) s1 += D[r * cols + c];
            const double mu = s1 / (double)count;
            for (int64_t r = 0; r < count; ++r) {
                const double d = D[r * cols + c] - mu;
                s2 += d * d;
            }
            mean[c] = (float)mu;
            sd[c] = (float)std::max(std::sqrt(s2 / (double)count), floor_sd);
        }
    };
    stats(X, kMlpIn, w + L.in_mean, w + L.in_std, 1e-6);
    stats(Y, kMlpOut, w + L.out_mean, w + L.out_std, 1e-12);
}

// Minibatch Adam with a cosine learning-rate schedule over `epochs` passes
// of `count` device samples (dX, dY); one shuffle per epoch (SplitMix64).
hipError_t mlp_train(float* w, const float* dX, const float* dY, int64_t count, int epochs, int batch, float lr,
                     uint64_t seed, float* loss_per_epoch, hipStream_t s)
{
    const MlpLayout L = mlp_layout();
    if (batch > count) batch = (int)count;
    const int64_t nb = count / batch;
    float* work = nullptr;
    const int64_t nbuf = mlp_buffer_floats(batch);
    const int64_t need = 3 * L.total + nbuf + (int64_t)batch * (kMlpIn + kMlpOut);
    hipError_t e = hipMalloc(&work, sizeof(float) * need);
    if (e != hipSuccess) return e;
    int64_t* didx = nullptr;
    double* dacc = nullptr;
    e = hipMalloc(&didx, sizeof(int64_t) * count);
    if (e == hipSuccess) e = hipMalloc(&dacc, sizeof(double));
    std::vector<int64_t> idx(count);
    for (int64_t i = 0; i < count; ++i) idx[i] = i;
    uint64_t state = seed ^ 0xA0761D6478BD642Full;
    auto next = [&state]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    float* g = work;
    float* m = g + L.total;
    float* v = m + L.total;
    MlpBuffers buf;
    mlp_buffer_bind(v + L.total, batch, &buf);
    float* bX = v + L.total + nbuf;
    float* bY = bX + (int64_t)batch * kMlpIn;
    if (e == hipSuccess) e = hipMemsetAsync(work, 0, sizeof(float) * 3 * L.total, s);
    int step = 0;
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
        for (int64_t i = count - 1; i > 0; --i) std::swap(idx[i], idx[(int64_t)(next() % (uint64_t)(i + 1))]);
        if ((e = hipMemcpyAsync(didx, idx.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice, s)) != hipSuccess) break;
        if ((e = hipMemsetAsync(dacc, 0, sizeof(double), s)) != hipSuccess) break;
        const float lr_e = lr * 0.5f * (1.f + std::cos(3.14159265358979f * (float)ep / (float)epochs));
        for (int64_t b = 0; b < nb && e == hipSuccess; ++b) {
            const int64_t* ib = didx + b * batch;
            if ((e = mlp_gather(batch, kMlpIn, dX, ib, bX, s)) != hipSuccess) break;
            if ((e = mlp_gather(batch, kMlpOut, dY, ib, bY, s)) != hipSuccess) break;
            if ((e = mlp_forward(w, batch, bX, buf, s)) != hipSuccess) break;
            if ((e = mlp_loss(w, batch, bY, buf, s)) != hipSuccess) break;
            if ((e = mlp_accumulate_loss(batch, buf, dacc, s)) != hipSuccess) break;
            if ((e = mlp_backward(w, batch, buf, g, s)) != hipSuccess) break;
            e = mlp_adam((int)L.num_params, w + L.first_param, g + L.first_param, m + L.first_param,
                         v + L.first_param, lr_e, ++step, s);
        }
        if (e != hipSuccess) break;
        double acc = 0.0;
        if ((e = hipMemcpyAsync(&acc, dacc, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(acc / ((double)nb * batch * kMlpOut));
    }
    hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    (void)hipFree(work);
    if (didx) (void)hipFree(didx);
    if (dacc) (void)hipFree(dacc);
    return e;
}

}  // namespace rf
