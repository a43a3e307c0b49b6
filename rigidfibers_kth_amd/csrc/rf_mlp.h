// rf_mlp.h -- the learned initial guess (internal): per-fiber features, a
// 113-128-128-128-15 tanh MLP in FP32 and its training kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

constexpr int kMlpNeighbours = 16;
constexpr int kMlpIn = 7 * kMlpNeighbours + 1;  // 16 x (rel. centre, orientation, proximity) + t.z
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 15;                     // 3M coefficients, M = 5
constexpr int kMlpTerms = 5;

// Offsets (floats) into the weight blob.
struct MlpLayout {
    int64_t in_mean, in_std, first_param, W[4], b[4], num_params, out_mean, out_std, total;
};
MlpLayout mlp_layout();

// Activation and gradient buffers for a batch of up to `rows` rows.
struct MlpBuffers {
    int rows = 0;
    float* Z = nullptr;        // [rows][kMlpIn] standardised input
    float* H[3] = {nullptr};   // [rows][kMlpHidden]
    float* Yhat = nullptr;     // [rows][kMlpOut]
    float* D[4] = {nullptr};   // deltas: D[0..2] [rows][kMlpHidden], D[3] [rows][kMlpOut]
    float* partial = nullptr;  // per-block partial losses
};
int64_t mlp_buffer_floats(int rows);
void mlp_buffer_bind(float* base, int rows, MlpBuffers* b);
int64_t mlp_loss_blocks(int B);

// X[i][kMlpIn] from the first 16 entries of each neighbour list (ldn >= 16).
hipError_t mlp_features(int N, int ldn, const double* pos, const double* orient, const int* nbr,
                        const double* nbr_dist, float* X, hipStream_t s);
// Y[i][3M]: coefficients of fiber i in its local frame.
hipError_t mlp_targets(int N, int M, const double* orient, const double* x, float* Y, hipStream_t s);
hipError_t mlp_forward(const float* w, int B, const float* X, MlpBuffers& buf, hipStream_t s);
// buf.D[3] = d loss / d Yhat for the mean squared error on standardised
// targets; buf.partial[blocks] = partial sums of squared errors.
hipError_t mlp_loss(const float* w, int B, const float* Y, MlpBuffers& buf, hipStream_t s);
hipError_t mlp_backward(const float* w, int B, MlpBuffers& buf, float* g, hipStream_t s);
// x (global, fiber-major 3M) from the de-standardised local prediction.
hipError_t mlp_to_global(int N, int M, const double* orient, const float* w, const float* Yhat, double* x,
                         hipStream_t s);
hipError_t mlp_adam(int n, float* w, const float* g, float* m, float* v, float lr, int step, hipStream_t s);
hipError_t mlp_gather(int B, int W, const float* src, const int64_t* idx, float* dst, hipStream_t s);
// acc[0] += the partial losses of the last mlp_loss.
hipError_t mlp_accumulate_loss(int B, const MlpBuffers& buf, double* acc, hipStream_t s);

// Host-side: Glorot-uniform weights (SplitMix64), standardisation from data.
void mlp_init_weights(uint64_t seed, float* w);
void mlp_set_standardisation(const float* X, const float* Y, int64_t count, float* w);
// Minibatch Adam (cosine schedule) on device samples; w is the device blob.
hipError_t mlp_train(float* w, const float* dX, const float* dY, int64_t count, int epochs, int batch, float lr,
                     uint64_t seed, float* loss_per_epoch, hipStream_t s);

}  // namespace rf
