// FP32 kernels of the row-size predictor (see mlp.h): forward, bin
// prediction, per-row backward deltas, ordered gradient reduction, Adam.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace sdl {

namespace {

unsigned grid_of(int64_t n, int b = 256) { return (unsigned) std::max<int64_t>(1, (n + b - 1) / b); }

// Parameters staged in LDS by every workgroup.
__device__ __forceinline__ void stage_params(float *sp, const float *__restrict__ P)
{
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) sp[i] = P[i];
    __syncthreads();
}

// pre[j] = (W1 x)_j + b1_j, h = relu(pre), z = W2 h + b2 (sums in index order)
__device__ __forceinline__ void mlp_row(const float *sp, const float *__restrict__ x, float pre[MLP_HID],
                                        float h[MLP_HID], float z[MLP_OUT])
{
    float xin[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) xin[i] = x[i];
#pragma unroll
    for (int j = 0; j < MLP_HID; j++)
    {
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) a = a + sp[j * MLP_IN + i] * xin[i];
        a = a + sp[MLP_OFF_B1 + j];
        pre[j] = a;
        h[j] = a > 0.0f ? a : 0.0f;
    }
#pragma unroll
    for (int c = 0; c < MLP_OUT; c++)
    {
        float a = 0.0f;
#pragma unroll
        for (int j = 0; j < MLP_HID; j++) a = a + sp[MLP_OFF_W2 + c * MLP_HID + j] * h[j];
        z[c] = a + sp[MLP_OFF_B2 + c];
    }
}

__device__ __forceinline__ void softmax6(const float z[MLP_OUT], float p[MLP_OUT], float *lse)
{
    float mx = z[0];
#pragma unroll
    for (int c = 1; c < MLP_OUT; c++) mx = fmaxf(mx, z[c]);
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < MLP_OUT; c++)
    {
        p[c] = expf(z[c] - mx);
        s = s + p[c];
    }
#pragma unroll
    for (int c = 0; c < MLP_OUT; c++) p[c] = p[c] / s;
    *lse = mx + logf(s);
}

__global__ void __launch_bounds__(256)
    k_mlp_forward(int64_t n, const float *__restrict__ X, const float *__restrict__ P,
                  float *__restrict__ Z)
{
    __shared__ float sp[MLP_NPARAM];
    stage_params(sp, P);
    int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float pre[MLP_HID], h[MLP_HID], z[MLP_OUT];
    mlp_row(sp, X + r * MLP_IN, pre, h, z);
#pragma unroll
    for (int c = 0; c < MLP_OUT; c++) Z[r * MLP_OUT + c] = z[c];
}

__global__ void __launch_bounds__(256)
    k_mlp_predict(int64_t n, const float *__restrict__ X, const float *__restrict__ P, float tau,
                  unsigned char *__restrict__ bins)
{
    __shared__ float sp[MLP_NPARAM];
    stage_params(sp, P);
    int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float pre[MLP_HID], h[MLP_HID], z[MLP_OUT], p[MLP_OUT], lse;
    mlp_row(sp, X + r * MLP_IN, pre, h, z);
    softmax6(z, p, &lse);
    float cum = 0.0f;
    int b = MLP_OUT - 1;
    for (int c = 0; c < MLP_OUT; c++)
    {
        cum = cum + p[c];
        if (cum >= tau)
        {
            b = c;
            break;
        }
    }
    bins[r] = (unsigned char) b;
}

// Per-row deltas: H = h, DZ = (p - onehot(y)) / n, DP = (W2^T DZ) * relu'(pre),
// loss = lse - z[y].
__global__ void __launch_bounds__(256)
    k_mlp_rows(int64_t n, const float *__restrict__ X, const int *__restrict__ Y,
               const float *__restrict__ P, float *__restrict__ H, float *__restrict__ DZ,
               float *__restrict__ DP, float *__restrict__ loss)
{
    __shared__ float sp[MLP_NPARAM];
    stage_params(sp, P);
    int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float pre[MLP_HID], h[MLP_HID], z[MLP_OUT], p[MLP_OUT], lse;
    mlp_row(sp, X + r * MLP_IN, pre, h, z);
    softmax6(z, p, &lse);
    const int y = Y[r];
    const float inv_n = 1.0f / (float) n;
    float dz[MLP_OUT];
#pragma unroll
    for (int c = 0; c < MLP_OUT; c++)
    {
        dz[c] = (p[c] - (c == y ? 1.0f : 0.0f)) * inv_n;
        DZ[r * MLP_OUT + c] = dz[c];
    }
#pragma unroll
    for (int j = 0; j < MLP_HID; j++)
    {
        float a = 0.0f;
#pragma unroll
        for (int c = 0; c < MLP_OUT; c++) a = a + sp[MLP_OFF_W2 + c * MLP_HID + j] * dz[c];
        H[r * MLP_HID + j] = h[j];
        DP[r * MLP_HID + j] = pre[j] > 0.0f ? a : 0.0f;
    }
    float zy = z[0];
#pragma unroll
    for (int c = 1; c < MLP_OUT; c++)
        if (c == y) zy = z[c];
    loss[r] = lse - zy;
}

// Gradient reduction in two deterministic stages.  Stage 1: one workgroup per
// chunk of MLP_RCH rows (staged in LDS); each parameter's chunk partial sums
// the chunk's rows in ascending order.  Stage 2: one thread per parameter sums
// the chunk partials in ascending chunk order.
constexpr int MLP_RCH = 64;

__global__ void __launch_bounds__(256)
    k_mlp_reduce_chunks(int64_t n, const float *__restrict__ X, const float *__restrict__ H,
                        const float *__restrict__ DZ, const float *__restrict__ DP,
                        float *__restrict__ part)
{
    __shared__ float sx[MLP_RCH][MLP_IN], sh[MLP_RCH][MLP_HID], sz[MLP_RCH][MLP_OUT],
        sp[MLP_RCH][MLP_HID];
    const int64_t r0 = (int64_t) blockIdx.x * MLP_RCH;
    const int nr = (int) min((int64_t) MLP_RCH, n - r0);
    for (int i = threadIdx.x; i < nr * MLP_IN; i += blockDim.x)
        sx[i / MLP_IN][i % MLP_IN] = X[r0 * MLP_IN + i];
    for (int i = threadIdx.x; i < nr * MLP_HID; i += blockDim.x)
    {
        sh[i / MLP_HID][i % MLP_HID] = H[r0 * MLP_HID + i];
        sp[i / MLP_HID][i % MLP_HID] = DP[r0 * MLP_HID + i];
    }
    for (int i = threadIdx.x; i < nr * MLP_OUT; i += blockDim.x)
        sz[i / MLP_OUT][i % MLP_OUT] = DZ[r0 * MLP_OUT + i];
    __syncthreads();
    for (int p = threadIdx.x; p < MLP_NPARAM; p += blockDim.x)
    {
        float acc = 0.0f;
        if (p < MLP_OFF_B1)
        {
            const int j = p / MLP_IN, i = p % MLP_IN;
            for (int r = 0; r < nr; r++) acc = acc + sp[r][j] * sx[r][i];
        }
        else if (p < MLP_OFF_W2)
        {
            const int j = p - MLP_OFF_B1;
            for (int r = 0; r < nr; r++) acc = acc + sp[r][j];
        }
        else if (p < MLP_OFF_B2)
        {
            const int c = (p - MLP_OFF_W2) / MLP_HID, j = (p - MLP_OFF_W2) % MLP_HID;
            for (int r = 0; r < nr; r++) acc = acc + sz[r][c] * sh[r][j];
        }
        else
        {
            const int c = p - MLP_OFF_B2;
            for (int r = 0; r < nr; r++) acc = acc + sz[r][c];
        }
        part[(int64_t) blockIdx.x * MLP_NPARAM + p] = acc;
    }
}

__global__ void __launch_bounds__(256)
    k_mlp_reduce(int nchunks, const float *__restrict__ part, float *__restrict__ G)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float acc = 0.0f;
    for (int c = 0; c < nchunks; c++) acc = acc + part[(int64_t) c * MLP_NPARAM + p];
    G[p] = acc;
}

// Mean of loss[0..n): contiguous chunks per thread, then a fixed tree.
__global__ void __launch_bounds__(256) k_mean(int64_t n, const float *__restrict__ v, float *out)
{
    __shared__ float sh[256];
    const int64_t per = (n + 255) / 256;
    const int64_t a = threadIdx.x * per, b = std::min<int64_t>(n, a + per);
    float s = 0.0f;
    for (int64_t i = a; i < b; i++) s = s + v[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int) threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0] / (float) n;
}

__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M,
                           float *__restrict__ V, float lr, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = G[p];
    const float m = b1 * M[p] + (1.0f - b1) * g;
    const float v = b2 * V[p] + (1.0f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    // bias-corrected: c1 = 1 - b1^t, c2 = 1 - b2^t
    P[p] = P[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_bins_of_counts(int64_t n, const int64_t *__restrict__ rp, int *__restrict__ Y)
{
    int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) Y[r] = mlp_bin_of(rp[r + 1] - rp[r]);
}

__global__ void k_gather_batch(int64_t nb, const int *__restrict__ perm, const float *__restrict__ X,
                               const int *__restrict__ Y, float *__restrict__ Xb,
                               int *__restrict__ Yb)
{
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int64_t r = perm[i];
#pragma unroll
    for (int f = 0; f < MLP_IN; f++) Xb[i * MLP_IN + f] = X[r * MLP_IN + f];
    Yb[i] = Y[r];
}

// splitmix64 stream for the weight init and the batch shuffles
uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

} // namespace

const RowPredictor *g_row_predictor = nullptr;

void MlpScratch::reserve(int64_t n)
{
    if (n <= cap) return;
    H.alloc((size_t) n * MLP_HID);
    DZ.alloc((size_t) n * MLP_OUT);
    DP.alloc((size_t) n * MLP_HID);
    loss.alloc((size_t) n + 1);
    part.alloc((size_t) ((n + MLP_RCH - 1) / MLP_RCH) * MLP_NPARAM);
    cap = n;
}

void mlp_init(RowPredictor &R, uint64_t seed)
{
    // uniform(-a, a) with a = sqrt(6 / (fan_in + fan_out)) per layer, zero biases
    std::vector<float> h(MLP_NPARAM, 0.0f);
    uint64_t s = seed;
    auto uni = [&](float a) { return (float) (((splitmix(s) >> 11) * 0x1.0p-53) * 2.0 - 1.0) * a; };
    const float a1 = std::sqrt(6.0f / (MLP_IN + MLP_HID)), a2 = std::sqrt(6.0f / (MLP_HID + MLP_OUT));
    for (int i = 0; i < MLP_OFF_B1; i++) h[i] = uni(a1);
    for (int i = MLP_OFF_W2; i < MLP_OFF_B2; i++) h[i] = uni(a2);
    R.params = upload(h.data(), MLP_NPARAM);
    std::vector<float> z(MLP_NPARAM, 0.0f);
    R.m = upload(z.data(), MLP_NPARAM);
    R.v = upload(z.data(), MLP_NPARAM);
    R.grad = upload(z.data(), MLP_NPARAM);
    R.step = 0;
    sync();
}

void mlp_forward(const RowPredictor &R, const float *X, int64_t n, float *Z, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid_of(n)), dim3(256), 0, s, n, X, R.params.p, Z);
    SDL_HIP(hipGetLastError());
}

void mlp_predict(const RowPredictor &R, const float *X, int64_t n, unsigned char *bins,
                 hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_predict, dim3(grid_of(n)), dim3(256), 0, s, n, X, R.params.p, R.tau,
                       bins);
    SDL_HIP(hipGetLastError());
}

// The launches of mlp_gradients; the mean loss goes to loss_out (device).
void mlp_gradients_launch(const RowPredictor &R, const float *X, const int *Y, int64_t n,
                          float *G, MlpScratch &S, hipStream_t s, float *loss_out)
{
    S.reserve(n);
    hipLaunchKernelGGL(k_mlp_rows, dim3(grid_of(n)), dim3(256), 0, s, n, X, Y, R.params.p, S.H.p,
                       S.DZ.p, S.DP.p, S.loss.p);
    SDL_HIP(hipGetLastError());
    const int nchunks = (int) ((n + MLP_RCH - 1) / MLP_RCH);
    hipLaunchKernelGGL(k_mlp_reduce_chunks, dim3(nchunks), dim3(256), 0, s, n, X, S.H.p, S.DZ.p,
                       S.DP.p, S.part.p);
    SDL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3(grid_of(MLP_NPARAM)), dim3(256), 0, s, nchunks,
                       S.part.p, G);
    SDL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mean, dim3(1), dim3(256), 0, s, n, S.loss.p, loss_out);
    SDL_HIP(hipGetLastError());
}

float mlp_gradients(const RowPredictor &R, const float *X, const int *Y, int64_t n, float *G,
                    MlpScratch &S, hipStream_t s)
{
    SDL_REQUIRE(n > 0, "mlp_gradients: empty batch");
    S.reserve(n);
    mlp_gradients_launch(R, X, Y, n, G, S, s, S.loss.p + n);
    float loss = 0.0f;
    SDL_HIP(hipMemcpyAsync(&loss, S.loss.p + n, sizeof(float), hipMemcpyDeviceToHost, s));
    SDL_HIP(hipStreamSynchronize(s));
    return loss;
}

void mlp_adam(RowPredictor &R, const float *G, float lr, hipStream_t s)
{
    R.step++;
    const float c1 = 1.0f - std::pow(0.9f, (float) R.step);
    const float c2 = 1.0f - std::pow(0.999f, (float) R.step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(grid_of(MLP_NPARAM)), dim3(256), 0, s, R.params.p, G,
                       R.m.p, R.v.p, lr, c1, c2);
    SDL_HIP(hipGetLastError());
}

void mlp_bins_of_counts(const int64_t *rowptr, int64_t n, int *Y, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bins_of_counts, dim3(grid_of(n)), dim3(256), 0, s, n, rowptr, Y);
    SDL_HIP(hipGetLastError());
}

float mlp_train(RowPredictor &R, MlpScratch &S, const float *X, const int *Y, int64_t n,
                int epochs, int batch, float lr, uint64_t seed, hipStream_t s)
{
    SDL_REQUIRE(n > 0 && batch > 0 && epochs > 0, "mlp_train: empty data or batch");
    SDL_REQUIRE(n <= INT32_MAX, "mlp_train: too many rows");
    const int64_t B = std::min<int64_t>(batch, n);
    DBuf<int> dperm(n);
    DBuf<float> Xb((size_t) B * MLP_IN);
    DBuf<int> Yb(B);
    std::vector<int> perm(n);
    const int64_t nb = n / B;  // full batches per epoch
    DBuf<float> bloss(nb);
    std::vector<float> hloss(nb);
    uint64_t st = seed;
    float last = 0.0f;
    for (int e = 0; e < epochs; e++)
    {
        for (int64_t i = 0; i < n; i++) perm[i] = (int) i;
        for (int64_t i = n - 1; i > 0; i--)  // Fisher-Yates
        {
            const int64_t j = (int64_t) (splitmix(st) % (uint64_t) (i + 1));
            std::swap(perm[i], perm[j]);
        }
        SDL_HIP(hipMemcpyAsync(dperm.p, perm.data(), (size_t) n * sizeof(int),
                               hipMemcpyHostToDevice, s));
        for (int64_t b = 0; b < nb; b++)  // full batches only; no host sync inside the epoch
        {
            hipLaunchKernelGGL(k_gather_batch, dim3(grid_of(B)), dim3(256), 0, s, B,
                               dperm.p + b * B, X, Y, Xb.p, Yb.p);
            SDL_HIP(hipGetLastError());
            mlp_gradients_launch(R, Xb.p, Yb.p, B, R.grad.p, S, s, bloss.p + b);
            mlp_adam(R, R.grad.p, lr, s);
        }
        // the batch losses in order (the copy also keeps perm alive until the
        // epoch's work is done)
        SDL_HIP(hipMemcpyAsync(hloss.data(), bloss.p, (size_t) nb * sizeof(float),
                               hipMemcpyDeviceToHost, s));
        SDL_HIP(hipStreamSynchronize(s));
        double sum = 0.0;
        for (int64_t b = 0; b < nb; b++) sum += hloss[b];
        last = (float) (sum / std::max<int64_t>(nb, 1));
    }
    SDL_HIP(hipStreamSynchronize(s));
    return last;
}

} // namespace sdl
