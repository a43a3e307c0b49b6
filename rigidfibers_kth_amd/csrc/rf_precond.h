// rf_precond.h -- fiber neighbourhoods and the neighbour-coupled block-Jacobi
// preconditioner (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rf {

constexpr int kMaxNeighbours = 16;    // k nearest fibers kept per fiber
constexpr int kMaxRasNeighbours = 6;  // neighbours per preconditioner cluster

// nbr[i*K + r], r < K: the K fibers nearest to fiber i by segment-segment
// distance (ascending, ties by index); nbr_dist the distances.
hipError_t launch_knn(int N, int K, const double* pos, const double* orient, int* nbr, double* nbr_dist,
                      hipStream_t s);
size_t ras_setup_lds(int K, int bs);
// B[i]: the bs rows of (A_{C_i C_i})^{-1} that belong to fiber i, C_i = {i} +
// nbr[i*ldn .. i*ldn+K); bs = 3M.  *info = i+1 if a cluster matrix is singular.
hipError_t launch_ras_setup(int N, int K, int ldn, int bs, const int* nbr, const void* A, int elem_bytes,
                            int64_t lda, double* B, int* info, hipStream_t s);
// z = M^{-1} r
hipError_t launch_ras_apply(int N, int K, int ldn, int bs, const int* nbr, const double* B, const double* r,
                            double* z, hipStream_t s);

}  // namespace rf
