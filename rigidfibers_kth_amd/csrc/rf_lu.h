// rf_lu.h -- dense LU with partial pivoting (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

// In-place LU of the n x n column-major matrix (double if elem_bytes == 8,
// float otherwise); ipiv[k] = row swapped with k (0-based); info[0] = k+1 if
// U(k,k) == 0.  use_mfma selects the matrix-core trailing update; panel
// width nb <= 128 for float, <= 64 for double.  ws:
// lu_workspace_bytes() of device memory whose last word (the column
// kernels' counter) is zero on entry and is left zero.
// block_pivot = 1 searches pivots inside each panel's diagonal block only
// and sets info[1] = 1 when that differs from partial pivoting (a multiplier
// |l| > 1 or a zero pivot); the matrix is then overwritten with factors that
// must be discarded, and the caller refactors the original with
// block_pivot = 0 (pivot search over whole columns).  info[0..1] are zero on
// entry.
size_t lu_workspace_bytes();
hipError_t lu_factor(void* A, int elem_bytes, int64_t lda, int n, int nb, int* ipiv, int* info, int use_mfma,
                     int block_pivot, void* ws, hipStream_t s);
// x <- U^-1 L^-1 P x with FP64 vectors (substitution, two launches per
// 64-row block).
hipError_t lu_solve(const void* LU, int elem_bytes, int64_t lda, int n, const int* ipiv, double* x, hipStream_t s);

// The same solve with one launch per 64-row block and sweep, for the factors
// that precondition every GMRES iteration.  aux holds lu_solve_aux_bytes(n):
// the row permutation, filled by lu_prepare_solve once per factorisation;
// tmp is n doubles.
size_t lu_solve_aux_bytes(int n);
hipError_t lu_prepare_solve(const void* LU, int elem_bytes, int64_t lda, int n, const int* ipiv, void* aux,
                            hipStream_t s);
hipError_t lu_solve_blocks(const void* LU, int elem_bytes, int64_t lda, int n, const void* aux, double* x,
                           double* tmp, hipStream_t s);

// Lane-layout probe of the 16x16x4 matrix-core instructions (one wavefront).
hipError_t mfma_probe(int mode, double* out64, float* out32, hipStream_t s);

}  // namespace rf
