// rf_step.hip -- per-fiber kernels of the timestep on gfx950: the source
// vector of the velocity operator, the velocity epilogue and the
// second-order multistep advance.
#include "rf_internal.h"

namespace rf {
namespace {

// fext[(j,l,c)] = F_j/2 for l = 0, a_j^l for l = 1..M: the force expansion
// f_j = F_j/2 + sum_l a_j^l P_l (fortran/matrix_assembly.f90:123-157).
__global__ void build_fext_kernel(int N, int M, const double* __restrict__ force, const double* __restrict__ x,
                                  double* __restrict__ fext)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = 3 * (M + 1);
    if (idx >= N * per) return;
    const int j = idx / per, r = idx - j * per;
    const int l = r / 3, c = r - l * 3;
    fext[idx] = (l == 0) ? 0.5 * force[3 * j + c] : x[(size_t)j * 3 * M + (l - 1) * 3 + c];
}

// V_i = [1/2((d+e)F + (d-e)(t.F)t) + sh d y_i e_x + 1/2 y_V]/d,
// Omega_i = [sh d t_y (e_x - t_x t) + 3/2 (y_O - (t.y_O) t)]/d
// (fortran/coeffs_and_vels.f90:596-600,619-627, mubar = d); y6[6i..6i+5] =
// (y_V, y_O); sh is the rate of the background shear flow (sh y, 0, 0).
__global__ void velocity_epilogue_kernel(int N, double d, double e, double sh, const double* __restrict__ pos,
                                         const double* __restrict__ orient, const double* __restrict__ force,
                                         const double* __restrict__ y6, double* __restrict__ vel,
                                         double* __restrict__ rot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double t0 = orient[3 * i], t1 = orient[3 * i + 1], t2 = orient[3 * i + 2];
    const double F0 = force[3 * i], F1 = force[3 * i + 1], F2 = force[3 * i + 2];
    const double tF = t0 * F0 + t1 * F1 + t2 * F2;
    const double* y = y6 + 6 * i;
    const double inv = 1.0 / d;
    const double shd = sh * d;
    vel[3 * i + 0] = (0.5 * ((d + e) * F0 + (d - e) * tF * t0) + shd * pos[3 * i + 1] + 0.5 * y[0]) * inv;
    vel[3 * i + 1] = (0.5 * ((d + e) * F1 + (d - e) * tF * t1) + 0.5 * y[1]) * inv;
    vel[3 * i + 2] = (0.5 * ((d + e) * F2 + (d - e) * tF * t2) + 0.5 * y[2]) * inv;
    const double ty = t0 * y[3] + t1 * y[4] + t2 * y[5];
    const double st = -shd * t1 * t0;
    rot[3 * i + 0] = (st * t0 + shd * t1 + 1.5 * (y[3] - ty * t0)) * inv;
    rot[3 * i + 1] = (st * t1 + 1.5 * (y[4] - ty * t1)) * inv;
    rot[3 * i + 2] = (st * t2 + 1.5 * (y[5] - ty * t2)) * inv;
}

// Euler first step / 2nd-order multistep + normalisation
// (fortran/advect_fibers.f90:498-523).
__global__ void advance_kernel(int N, int first, double dt, const double* __restrict__ xp,
                               const double* __restrict__ xc, const double* __restrict__ tp,
                               const double* __restrict__ tc, const double* __restrict__ vo,
                               const double* __restrict__ vn, const double* __restrict__ ro,
                               const double* __restrict__ rn, double* __restrict__ xnext, double* __restrict__ tnext)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double tt[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int k = 3 * i + c;
        if (first) {
            xnext[k] = xc[k] + dt * vn[k];
            tt[c] = tc[k] + dt * rn[k];
        } else {
            xnext[k] = 4.0 / 3.0 * xc[k] - 1.0 / 3.0 * xp[k] + 2.0 / 3.0 * dt * (2.0 * vn[k] - vo[k]);
            tt[c] = 4.0 / 3.0 * tc[k] - 1.0 / 3.0 * tp[k] + 2.0 / 3.0 * dt * (2.0 * rn[k] - ro[k]);
        }
    }
    const double m = sqrt(tt[0] * tt[0] + tt[1] * tt[1] + tt[2] * tt[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) tnext[3 * i + c] = tt[c] / m;
}

__global__ void fill_kernel(double* __restrict__ p, double v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void axpby_kernel(int n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                             double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + b * y[i];
}

// x1 = u x, x2 = x - x1 with u_i in [0, 1) from SplitMix64 of the index
// ((z >> 11) 2^-53): two random summands of x for a linearity test.
__global__ void random_split_kernel(int n, const double* __restrict__ x, double* __restrict__ x1,
                                    double* __restrict__ x2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * 0x1.0p-53;
    const double a = u * x[i];
    x1[i] = a;
    x2[i] = x[i] - a;
}

}  // namespace

hipError_t launch_random_split(int n, const double* x, double* x1, double* x2, hipStream_t s)
{
    hipLaunchKernelGGL(random_split_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, x, x1, x2);
    return hipGetLastError();
}

hipError_t launch_build_fext(int N, int M, const double* force, const double* x, double* fext, hipStream_t s)
{
    const int tot = N * 3 * (M + 1);
    hipLaunchKernelGGL(build_fext_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, N, M, force, x, fext);
    return hipGetLastError();
}

hipError_t launch_velocity_epilogue(int N, double d, double e, double sh, const double* pos, const double* orient,
                                    const double* force, const double* y6, double* vel, double* rot, hipStream_t s)
{
    hipLaunchKernelGGL(velocity_epilogue_kernel, dim3((N + 127) / 128), dim3(128), 0, s, N, d, e, sh, pos, orient,
                       force, y6, vel, rot);
    return hipGetLastError();
}

hipError_t launch_advance(int N, int first, double dt, const double* x_prev, const double* x_cur, const double* t_prev,
                          const double* t_cur, const double* v_old, const double* v_new, const double* r_old,
                          const double* r_new, double* x_next, double* t_next, hipStream_t s)
{
    hipLaunchKernelGGL(advance_kernel, dim3((N + 127) / 128), dim3(128), 0, s, N, first, dt, x_prev, x_cur, t_prev,
                       t_cur, v_old, v_new, r_old, r_new, x_next, t_next);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, double v, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_axpby(int n, double a, const double* x, double b, const double* y, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(axpby_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, a, x, b, y, out);
    return hipGetLastError();
}

}  // namespace rf
