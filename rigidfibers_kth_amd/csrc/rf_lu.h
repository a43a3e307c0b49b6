// rf_lu.h -- dense LU with partial pivoting (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

// In-place LU of the n x n column-major matrix (double if elem_bytes == 8,
// float otherwise); ipiv[k] = row swapped with k (0-based); info[0] = k+1 if
// U(k,k) == 0.  use_mfma selects the matrix-core trailing update; panel
// width nb <= 128 for float, <= 64 for double.  ws:
// lu_workspace_bytes(n, update_fp16) of device memory, zero on entry (the
// column kernels' counter in it is left zero).
// update_fp16 = 1 (float, use_mfma): trailing updates with fp16-rounded
// operands and FP32 accumulation, for factors that precondition; the
// workspace must then have been sized with update_fp16 = 1.
// block_pivot = 1 searches each panel's pivots inside its diagonal block
// first and checks the result against partial pivoting (every multiplier
// |l| <= 1, no zero pivot); a panel that fails is restored and factorised
// column by column (block_pivot = 0: every panel that way), so the pivots
// are partial pivoting's either way.  *redone (may be null) receives the
// number of panels redone.  Block pivoting synchronises the stream once per
// panel, reading the check through host_flag (two ints of coherent pinned
// host memory: the flag and the copy's sequence word, which the host polls)
// when it is not null.  info[0..1] are zero on entry.
size_t lu_workspace_bytes(int n, int update_fp16);
hipError_t lu_factor(void* A, int elem_bytes, int64_t lda, int n, int nb, int* ipiv, int* info, int use_mfma,
                     int block_pivot, int update_fp16, void* ws, int* redone, int* host_flag, hipStream_t s);
// x <- U^-1 L^-1 P x with FP64 vectors, one launch per 64-row block and
// sweep (FP32 or FP64 factors).  aux holds
// lu_solve_aux_bytes(n, elem_bytes): the row permutation and the inverses of
// the 64x64 diagonal blocks of L and U, filled by lu_prepare_solve once per
// factorisation; tmp is n doubles.
size_t lu_solve_aux_bytes(int n, int elem_bytes);
hipError_t lu_prepare_solve(const void* LU, int elem_bytes, int64_t lda, int n, const int* ipiv, void* aux,
                            hipStream_t s);
hipError_t lu_solve_blocks(const void* LU, int elem_bytes, int64_t lda, int n, const void* aux, double* x,
                           double* tmp, hipStream_t s);

// Lane-layout probe of the 16x16x4 matrix-core instructions (one wavefront).
hipError_t mfma_probe(int mode, double* out64, float* out32, hipStream_t s);

}  // namespace rf
