// Learned row binning: a 16 -> 32 -> 6 FP32 MLP that predicts, from cheap
// structural features of an output row of C = kron(I, L) A, which size bin
// the row's nnz falls in, so the symbolic phase can size each row's hash
// table without the exact counting pass.  Training, inference and the
// gradient computation are the library's own HIP kernels.
//
// Parameters (742 floats, one array): W1[32][16] | b1[32] | W2[6][32] | b2[6].
//   h = relu(W1 x + b1), z = W2 h + b2, p = softmax(z).
// Loss: mean cross-entropy over a batch.  Gradients are summed over the batch
// rows in row order by one thread per parameter (deterministic).
#pragma once

#include "sdl_internal.h"

namespace sdl {

constexpr int MLP_IN = 16;
constexpr int MLP_HID = 32;
constexpr int MLP_OUT = 6;
constexpr int MLP_NPARAM = MLP_HID * MLP_IN + MLP_HID + MLP_OUT * MLP_HID + MLP_OUT;  // 742
constexpr int MLP_OFF_B1 = MLP_HID * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_OUT * MLP_HID;

// Size bins of a row's nnz: bin b holds nnz <= MLP_BIN_MAX[b] (b < 5); bin 5
// is everything larger.  A predicted bin b < 5 sizes the row's hash table at
// 2 * MLP_BIN_MAX[b] slots.
constexpr int MLP_NBIN_BOUNDED = 5;
__host__ __device__ inline int mlp_bin_max(int b)
{
    return b == 0 ? 32 : b == 1 ? 128 : b == 2 ? 512 : b == 3 ? 2048 : 4096;
}
__host__ __device__ inline int mlp_bin_of(long long nnz)
{
    for (int b = 0; b < MLP_NBIN_BOUNDED; b++)
        if (nnz <= mlp_bin_max(b)) return b;
    return MLP_NBIN_BOUNDED;
}

struct RowPredictor
{
    DBuf<float> params;   // MLP_NPARAM
    DBuf<float> m, v;     // Adam moments
    DBuf<float> grad;     // last gradient
    int step = 0;
    float tau = 0.95f;    // decision: smallest bin whose cumulative probability >= tau
};

// Device-side per-batch scratch for the gradient kernels.
struct MlpScratch
{
    DBuf<float> H, DZ, DP, loss;   // n*32, n*6, n*32, n
    DBuf<float> part;              // per-chunk gradient partials: ceil(n/64) * 742
    int64_t cap = 0;
    void reserve(int64_t n);
};

void mlp_init(RowPredictor &R, uint64_t seed);
// Logits of n rows (Z[n][6]).
void mlp_forward(const RowPredictor &R, const float *X, int64_t n, float *Z, hipStream_t s);
// Predicted bins (tau rule) of n rows.
void mlp_predict(const RowPredictor &R, const float *X, int64_t n, unsigned char *bins,
                 hipStream_t s);
// Mean cross-entropy and its gradient (G[MLP_NPARAM], device) over n rows
// with labels Y (bins, int32).  Returns the loss (synchronizes).
float mlp_gradients(const RowPredictor &R, const float *X, const int *Y, int64_t n, float *G,
                    MlpScratch &S, hipStream_t s);
// One Adam step on R.params with gradient G.
void mlp_adam(RowPredictor &R, const float *G, float lr, hipStream_t s);
// Labels from exact row counts: Y[i] = bin of cnt[i].
void mlp_bins_of_counts(const int64_t *rowptr, int64_t n, int *Y, hipStream_t s);
// Adam over shuffled mini-batches for `epochs` passes; returns the mean batch
// loss of the last epoch.
float mlp_train(RowPredictor &R, MlpScratch &S, const float *X, const int *Y, int64_t n,
                int epochs, int batch, float lr, uint64_t seed, hipStream_t s);

// Predictor used by the reference-compatible sparse product allocs (null:
// exact symbolic passes).
extern const RowPredictor *g_row_predictor;

} // namespace sdl
