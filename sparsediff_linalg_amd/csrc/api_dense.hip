// Drop-in dense helpers (reference: include/sparsediff_linalg/dense.h,
// src/support/mini_numpy.c:22-105) on the GPU.  The ABI hands over host
// arrays, so each call uploads its inputs on the library stream, runs one
// kernel and copies the result back; the device entry points in sdl_device.h
// take HBM-resident arrays.  The two Kronecker products keep the reference's
// arithmetic: every output is a left-to-right sum from +0.0 of individually
// rounded products (no fused multiply-add), in the reference's index order.
#include "api_common.h"
#include "../../include/sdl_device.h"

#include <algorithm>
#include <climits>

extern "C" {
#include "sparsediff_linalg/dense.h"
}

namespace sdl {

namespace {

unsigned dense_grid(int64_t n) { return (unsigned) std::max<int64_t>(1, (n + 255) / 256); }

// result[i * repeats + r] = a[i]  (mini_numpy.c:22-32)
__global__ void k_repeat(int64_t total, int repeats, const double *__restrict__ a,
                         double *__restrict__ out)
{
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) out[t] = a[t / repeats];
}

// result[t * len + i] = a[i]  (mini_numpy.c:34-40)
__global__ void k_tile(int64_t total, int len, const double *__restrict__ a,
                       double *__restrict__ out)
{
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) out[t] = a[t % len];
}

__global__ void k_fill_value(int64_t n, double value, double *__restrict__ out)
{
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = value;
}

// AT[j*m + i] = A[i*n + j] for row-major m x n A: 32 x 32 tiles through LDS,
// reads and writes both coalesced (mini_numpy.c:50-72).
constexpr int TT = 32;
__global__ __launch_bounds__(TT * 8) void k_transpose_dense(int m, int n,
                                                          const double *__restrict__ A,
                                                          double *__restrict__ AT)
{
    __shared__ double tile[TT][TT + 1];
    const int ntj = (n + TT - 1) / TT;
    const int i0 = (int) (blockIdx.x / ntj) * TT, j0 = (int) (blockIdx.x % ntj) * TT;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 32 x 8
    for (int r = ty; r < TT; r += 8)
    {
        const int i = i0 + r, j = j0 + tx;
        if (i < m && j < n) tile[r][tx] = A[(int64_t) i * n + j];
    }
    __syncthreads();
    for (int r = ty; r < TT; r += 8)
    {
        const int j = j0 + r, i = i0 + tx;
        if (i < m && j < n) AT[(int64_t) j * m + i] = tile[tx][r];
    }
}

// Ordered dot-product tiles: out[p * op + q * oq] = sum over s ascending of
// P(p, s) * Q(q, s), from +0.0, with P(p, s) = P[p * pp + s * psd] and
// Q(q, s) = Q[q * qq + s * qsd].  A DT x DT output tile per workgroup of TS x TS
// threads, RB x RB outputs per thread (p = tile + tp + TS i, q = tile + tq +
// TS j); slices of DS reduction steps of P and Q go through LDS.  Each output
// keeps its own accumulator and adds its products in ascending s, so the sum is
// the reference's whatever the tiling (no fused multiply-add: built with
// -ffp-contract=off).  Few outputs and long sums take the smaller tiles, so
// the work still spreads over the CUs, and deeper slices, so each global-load
// round trip feeds more steps.
template <int RB, int TS, int DS>
__global__ __launch_bounds__(TS * TS) void k_ordered_dots(int np, int nq, int ns,
                                                     const double *__restrict__ P, int64_t pp,
                                                     int64_t psd, const double *__restrict__ Q,
                                                     int64_t qq, int64_t qsd,
                                                     double *__restrict__ out, int64_t op,
                                                     int64_t oq)
{
    constexpr int DT = TS * RB, NT = TS * TS;
    __shared__ double Ps[DS][DT + 1], Qs[DS][DT + 1];
    const int ntp = (np + DT - 1) / DT;
    const int bp = (int) (blockIdx.x % ntp), bq = (int) (blockIdx.x / ntp);
    const int tp = threadIdx.x % TS, tq = threadIdx.x / TS;
    double acc[RB][RB];
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < RB; j++) acc[i][j] = 0.0;
    for (int s0 = 0; s0 < ns; s0 += DS)
    {
        // consecutive threads walk the unit-stride index of each operand
        for (int idx = threadIdx.x; idx < DS * DT; idx += NT)
        {
            const int s = psd == 1 ? idx % DS : idx / DT, c = psd == 1 ? idx / DS : idx % DT;
            const int gp = bp * DT + c, gs = s0 + s;
            Ps[s][c] = (gp < np && gs < ns) ? P[gp * pp + gs * psd] : 0.0;
        }
        for (int idx = threadIdx.x; idx < DS * DT; idx += NT)
        {
            const int s = qsd == 1 ? idx % DS : idx / DT, c = qsd == 1 ? idx / DS : idx % DT;
            const int gq = bq * DT + c, gs = s0 + s;
            Qs[s][c] = (gq < nq && gs < ns) ? Q[gq * qq + gs * qsd] : 0.0;
        }
        __syncthreads();
        const int lim = ns - s0 < DS ? ns - s0 : DS;
#pragma unroll 4
        for (int s = 0; s < lim; s++)
        {
            double a[RB], b[RB];
#pragma unroll
            for (int i = 0; i < RB; i++) a[i] = Ps[s][tp + TS * i];
#pragma unroll
            for (int j = 0; j < RB; j++) b[j] = Qs[s][tq + TS * j];
#pragma unroll
            for (int i = 0; i < RB; i++)
#pragma unroll
                for (int j = 0; j < RB; j++)
                {
                    const double prod = a[i] * b[j];
                    acc[i][j] = acc[i][j] + prod;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < RB; i++)
    {
        const int p = bp * DT + tp + TS * i;
#pragma unroll
        for (int j = 0; j < RB; j++)
        {
            const int q = bq * DT + tq + TS * j;
            if (p < np && q < nq) out[p * op + q * oq] = acc[i][j];
        }
    }
}

void ordered_dots(int np, int nq, int ns, const double *P, int64_t pp, int64_t psd,
                  const double *Q, int64_t qq, int64_t qsd, double *out, int64_t op, int64_t oq,
                  hipStream_t st)
{
    if (np <= 0 || nq <= 0) return;
    auto tiles = [&](int dt) { return ((np + dt - 1) / dt) * (int64_t) ((nq + dt - 1) / dt); };
    constexpr int64_t enough = 512;  // two workgroups per CU
    if (tiles(64) >= enough)
        hipLaunchKernelGGL((k_ordered_dots<4, 16, 16>), dim3((unsigned) tiles(64)), dim3(256), 0, st,
                           np, nq, ns, P, pp, psd, Q, qq, qsd, out, op, oq);
    else if (tiles(32) >= enough)
        hipLaunchKernelGGL((k_ordered_dots<2, 16, 32>), dim3((unsigned) tiles(32)), dim3(256), 0, st,
                           np, nq, ns, P, pp, psd, Q, qq, qsd, out, op, oq);
    else if (tiles(16) >= enough)
        hipLaunchKernelGGL((k_ordered_dots<1, 16, 64>), dim3((unsigned) tiles(16)), dim3(256), 0, st,
                           np, nq, ns, P, pp, psd, Q, qq, qsd, out, op, oq);
    else
        hipLaunchKernelGGL((k_ordered_dots<1, 8, 256>), dim3((unsigned) tiles(8)), dim3(64), 0, st,
                           np, nq, ns, P, pp, psd, Q, qq, qsd, out, op, oq);
    SDL_HIP(hipGetLastError());
}

// v = (Y kron I_m) w: v[row + j*m] = sum_col Y[j + col*k] * w[row + col*m]
// (mini_numpy.c:74-89): p = row (np = m), q = j (nq = k), s = col (ns = n).
void y_kron_i(int m, int k, int n, const double *dY, const double *dw, double *dv, hipStream_t st)
{
    ordered_dots(m, k, n, dw, 1, m, dY, 1, k, dv, 1, m, st);
}

// v = (I_n kron X^T) w: v[j + col*k] = sum_row X[row + j*m] * w[row + col*m]
// (mini_numpy.c:91-105): p = j (np = k), q = col (nq = n), s = row (ns = m).
void i_kron_xt(int m, int k, int n, const double *dX, const double *dw, double *dv, hipStream_t st)
{
    ordered_dots(k, n, m, dX, m, 1, dw, m, 1, dv, 1, k, st);
}

int64_t checked_len(const char *fn, int64_t a, int64_t b)
{
    if (a < 0 || b < 0) api_exit(fn, "negative size");
    return a * b;
}

} // namespace

} // namespace sdl

using namespace sdl;

extern "C" {
#pragma GCC visibility push(default)

void repeat(double *result, const double *a, int len, int repeats)
{
    const int64_t total = checked_len("repeat", len, repeats);
    if (total == 0) return;
    DBuf<double> da = upload(a, len), dout(total);
    hipLaunchKernelGGL(k_repeat, dim3(dense_grid(total)), dim3(256), 0, stream(), total, repeats,
                       da.p, dout.p);
    SDL_HIP(hipGetLastError());
    d2h(result, dout.p, (size_t) total * sizeof(double));
}

void tile_double(double *result, const double *a, int len, int tiles)
{
    const int64_t total = checked_len("tile_double", len, tiles);
    if (total == 0) return;
    DBuf<double> da = upload(a, len), dout(total);
    hipLaunchKernelGGL(k_tile, dim3(dense_grid(total)), dim3(256), 0, stream(), total, len, da.p,
                       dout.p);
    SDL_HIP(hipGetLastError());
    d2h(result, dout.p, (size_t) total * sizeof(double));
}

void scaled_ones(double *result, int size, double value)
{
    if (size <= 0) return;
    DBuf<double> dout(size);
    hipLaunchKernelGGL(k_fill_value, dim3(dense_grid(size)), dim3(256), 0, stream(),
                       (int64_t) size, value, dout.p);
    SDL_HIP(hipGetLastError());
    d2h(result, dout.p, (size_t) size * sizeof(double));
}

void A_transpose(double *AT, const double *A, int m, int n)
{
    const int64_t total = checked_len("A_transpose", m, n);
    if (total == 0) return;
    DBuf<double> dA = upload(A, total), dAT(total);
    const unsigned blocks = (unsigned) (((n + TT - 1) / TT) * (int64_t) ((m + TT - 1) / TT));
    hipLaunchKernelGGL(k_transpose_dense, dim3(blocks), dim3(TT * 8), 0, stream(), m, n, dA.p,
                       dAT.p);
    SDL_HIP(hipGetLastError());
    d2h(AT, dAT.p, (size_t) total * sizeof(double));
}

void Y_kron_I_vec(int m, int k, int n, const double *Y, const double *w, double *v)
{
    const int64_t nv = checked_len("Y_kron_I_vec", m, k);
    if (nv == 0) return;
    DBuf<double> dY = upload(Y, checked_len("Y_kron_I_vec", k, n));
    DBuf<double> dw = upload(w, checked_len("Y_kron_I_vec", m, n)), dv(nv);
    y_kron_i(m, k, n, dY.p, dw.p, dv.p, stream());
    d2h(v, dv.p, (size_t) nv * sizeof(double));
}

void I_kron_XT_vec(int m, int k, int n, const double *X, const double *w, double *v)
{
    const int64_t nv = checked_len("I_kron_XT_vec", k, n);
    if (nv == 0) return;
    DBuf<double> dX = upload(X, checked_len("I_kron_XT_vec", m, k));
    DBuf<double> dw = upload(w, checked_len("I_kron_XT_vec", m, n)), dv(nv);
    i_kron_xt(m, k, n, dX.p, dw.p, dv.p, stream());
    d2h(v, dv.p, (size_t) nv * sizeof(double));
}

// ---- device-resident forms ----
int sdl_y_kron_i_vec(int m, int k, int n, const double *dY, const double *dw, double *dv,
                     void *stream)
{
    if (m < 0 || k < 0 || n < 0 || (m > 0 && k > 0 && (!dv || (n > 0 && (!dY || !dw)))))
    {
        set_error("sdl_y_kron_i_vec: invalid arguments");
        return 1;
    }
    y_kron_i(m, k, n, dY, dw, dv, stream ? (hipStream_t) stream : sdl::stream());
    return 0;
}

int sdl_i_kron_xt_vec(int m, int k, int n, const double *dX, const double *dw, double *dv,
                      void *stream)
{
    if (m < 0 || k < 0 || n < 0 || (k > 0 && n > 0 && (!dv || (m > 0 && (!dX || !dw)))))
    {
        set_error("sdl_i_kron_xt_vec: invalid arguments");
        return 1;
    }
    i_kron_xt(m, k, n, dX, dw, dv, stream ? (hipStream_t) stream : sdl::stream());
    return 0;
}

#pragma GCC visibility pop
} // extern "C"
