// rf_internal.h -- declarations shared by the HIP translation units of
// librigidfibers.so.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "rigidfibers.h"

namespace rf {

constexpr int kMaxTerms = 8;   // M <= 8 (fortran/matrix_assembly.f90:53-55)
constexpr int kMaxQuad = 96;   // LQ <= 96 (NoQI <= 32)

// Tables shared by all kernels of one simulation, built on the host.
struct Tables {
    int M = 0, LQ = 0;
    double eps = 0.0;
    double d = 0.0, e = 2.0, D1 = 0.0;
    double Ek[kMaxTerms + 1] = {0}, gammak[kMaxTerms + 1] = {0};
    double pv[kMaxQuad] = {0};  // quadrature points
    double wv[kMaxQuad] = {0};  // weights
    // wp[k*LQ + p] = w_p * P_k(s_p), k = 0..M (P_0 = 1)
    double wp[(kMaxTerms + 1) * kMaxQuad] = {0};
};

void build_tables(int M, int noqi, double eps, Tables* t);

// Assembly of the system matrix A (column-major, lda), the per-source-split
// partial right-hand sides and the velocity operator W.
struct AssembleArgs {
    int N = 0, n = 0;
    int64_t lda = 0;        // rows of A storage
    int64_t ldw = 0;        // rows of W storage (6N)
    int analytic = 1;
    int rhs_z_only = 1;     // analytic RHS reads only F_z (matrix_assembly.f90:857-859)
    int splits = 1;         // source splits (partial rhs slices)
    int targets_per_block = 0;
    const double* pos = nullptr;
    const double* orient = nullptr;
    const double* force = nullptr;
    const double* wp = nullptr;   // device copy of Tables::wp
    const double* pv = nullptr;   // device copy of Tables::pv
    void* A = nullptr;            // double* (FP64) or float* (mixed)
    double* bpart = nullptr;      // [splits][n]
    void* W = nullptr;            // [3(M+1)N cols][ldw rows], double or float; nullptr: not written
    int mixed = 0;                // 1: A and W stored as float
    double near_c = 10.0;         // mixed: (pair, point) with c < near_c computed in FP64
};

hipError_t launch_assemble(const Tables& t, const AssembleArgs& a, hipStream_t s);

// Matrix-free FP64 operator (rf_operator.hip).  Modes:
//   kOpAx       y = A x                         (n doubles)
//   kOpResidual y = b - A x                     (n doubles)
//   kOpRhs      y = b                           (n doubles; x unused)
//   kOpVelocity y = the velocity sums (IF_0, IF_1) per fiber, 6N doubles, the
//               product of the velocity operator W with [F/2, a^1..a^M]
enum { kOpAx = 0, kOpResidual = 1, kOpRhs = 2, kOpVelocity = 3 };
struct OperatorPlan {
    int N = 0, M = 0, LQ = 0, analytic = 1;
    int T = 1, threads = 64, splits = 1, src_per_split = 1, src_stride = 4;
};
OperatorPlan plan_operator(int N, int M, int LQ, int analytic, int num_cus);
// doubles of device scratch one launch needs (source vectors + split partials)
size_t operator_scratch_doubles(const OperatorPlan& p);
hipError_t launch_operator(const OperatorPlan& p, const Tables& t, int mode, const double* pos, const double* orient,
                           const double* force, const double* x, const double* wp, const double* pv, double* scratch,
                           double* y, hipStream_t s);
// The output map of `mode` alone, from the split partials the last
// launch_operator left in `scratch`: every mode's kernel forms all IF_k
// (k = 0..M), so the velocity sums (mode kOpVelocity) of an evaluation whose
// source vectors equal the velocity's -- a residual at the same x, with no
// x/y force component or numeric inner integrals -- need no second pass.
hipError_t launch_operator_epilogue(const OperatorPlan& p, const Tables& t, int mode, const double* orient,
                                    const double* x, const double* scratch, double* y, hipStream_t s);

// Column-major GEMV y = A x with FP64 arithmetic; A is double or float.
// Writes partial sums to parts[splits][rows_padded] then reduces into y.
struct GemvPlan {
    int rows = 0, cols = 0;
    int64_t ld = 0;
    int elem_bytes = 8;
    int splits = 1;
    int row_blocks = 0;
};
GemvPlan plan_gemv(int rows, int cols, int64_t ld, int elem_bytes, int num_cus);
size_t gemv_parts_elems(const GemvPlan& p);
// mode: 0 -> y = A x ; 1 -> y = b - A x ; scale applied to A x first.
// ev0/ev1 (optional) are recorded around the partial-product kernel only.
// With x_nrm2 the product is A (x / sqrt(*x_nrm2)), the scaled vector formed
// as launch_scale_copy forms it and stored to x_out when that is not null.
hipError_t launch_gemv(const GemvPlan& p, const void* A, const double* x, double* parts, double* y,
                       const double* b, int mode, hipStream_t s, hipEvent_t ev0 = nullptr,
                       hipEvent_t ev1 = nullptr, const double* x_nrm2 = nullptr, double* x_out = nullptr);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// bytes (a multiple of 4, at most 64 KiB) into coherent pinned host memory
hipError_t launch_to_host(const void* src, void* host_dst, size_t bytes, hipStream_t s,
                          unsigned* host_seq_word = nullptr, unsigned seq = 0);
hipError_t launch_f64_to_f32(const double* src, float* dst, size_t n, hipStream_t s);

// Deterministic reductions / vector kernels (rf_linalg.hip).
hipError_t launch_reduce_splits(const double* parts, int splits, int64_t stride, int n, double* out,
                                hipStream_t s);
// out[j] = sum_i V[i + j*ldv] * w[i] for j < k, out[k] = sum_i w[i]^2.
// parts must hold kMaxDotBlocks*(k+1) doubles (multidot_blocks(n) <= kMaxDotBlocks); counter is a device int (zero on entry, reset on exit).
hipError_t launch_multidot(const double* V, int64_t ldv, int k, const double* w, int n, double* parts,
                           unsigned int* counter, double* out, hipStream_t s);
// A copy of `words` 32-bit words from device memory into coherent pinned host
// memory that ends with `seq` stored to *seq_word at system scope (the
// sequence-word protocol of launch_to_host); dst == nullptr: no copy.
struct HostCopy {
    const unsigned* src = nullptr;
    unsigned* dst = nullptr;
    int words = 0;
    unsigned* seq_word = nullptr;
    unsigned seq = 0;
};
// w -= V[:, :k] h, then out[0..k) = V^T w (with_dots) and out[nd] = |w|^2
// (nd = k with dots, else 0), bit-identical to the update followed by
// launch_multidot; one kernel for k <= 32.  With hc.dst the same kernel then
// copies hc to the host (after out is written), as launch_to_host would.
hipError_t launch_orth_dots(const double* V, int64_t ldv, int k, const double* h, double* w, int n, int with_dots,
                            double* parts, unsigned int* counter, double* out, hipStream_t s,
                            const HostCopy& hc = HostCopy());
hipError_t launch_scale_copy(const double* w, const double* scale_num, int scale_idx, int n, double* v,
                             hipStream_t s);
// xc = x + V[:, :k] y, y in host memory (passed as kernel arguments)
hipError_t launch_vy_plus_x(const double* V, int64_t ldv, int k, const double* y_host, const double* x, int n,
                            double* xc, hipStream_t s);
// dst[0:n] = src[0:n] on the device (a kernel: no runtime copy packet)
hipError_t launch_copy_vec(const double* src, double* dst, int n, hipStream_t s);
int multidot_blocks(int n);
// workgroups per dot kernel at most: partial buffers hold kMaxDotBlocks *
// (slots) doubles
constexpr int kMaxDotBlocks = 64;

// Velocity epilogue, advance (rf_step.hip).
hipError_t launch_build_fext(int N, int M, const double* force, const double* x, double* fext, hipStream_t s);
hipError_t launch_velocity_epilogue(int N, double d, double e, double sh, const double* pos, const double* orient,
                                    const double* force, const double* y6, double* vel, double* rot,
                                    hipStream_t s);
hipError_t launch_advance(int N, int first, double dt, const double* x_prev, const double* x_cur,
                          const double* t_prev, const double* t_cur, const double* v_old, const double* v_new,
                          const double* r_old, const double* r_new, double* x_next, double* t_next,
                          hipStream_t s);
hipError_t launch_fill(double* p, double v, size_t n, hipStream_t s);
hipError_t launch_axpby(int n, double a, const double* x, double b, const double* y, double* out,
                        hipStream_t s);
// x1 = u x, x2 = x - x1 with fixed pseudo-random u_i in [0, 1)
hipError_t launch_random_split(int n, const double* x, double* x1, double* x2, hipStream_t s);

// Thread-local error message.
void set_error(const std::string& msg);

}  // namespace rf
