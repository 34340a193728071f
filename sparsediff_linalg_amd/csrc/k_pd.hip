// Generic kernels of the permuted_dense / stacked_pd kinds (see pd_kernels.h).
//
// ggemm runs every dense-block product of a fill as one launch: one
// workgroup per 64x64 tile of an output block, looping over the tile's tasks
// in order; each task's operands are gathered (row lists, column maps, d
// scaling) into LDS in 16-deep slices and multiplied with
// v_mfma_f64_16x16x4_f64 (lane l: a = A[l%16][l/16], b = B[l/16][l%16],
// acc[v] = D[l/16 + 4v][l%16]); four wavefronts own 32x32 quadrants.
#include "pd_kernels.h"

#include <algorithm>

namespace sdl {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

unsigned grid_for(int64_t n, int b = 256) { return (unsigned) std::max<int64_t>(1, (n + b - 1) / b); }

__global__ void k_pull_sum(int n, const int *__restrict__ seg, const int *__restrict__ idx,
                           const double *__restrict__ src, double *__restrict__ out)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double acc = 0.0;
    for (int s = seg[e]; s < seg[e + 1]; s++) acc = acc + src[idx[s]];
    out[e] = acc;
}

// one wavefront per output list, lanes over the dense dimension
__global__ void __launch_bounds__(256)
    k_sdmm(int NL, int NC, const int *__restrict__ lp, const int *__restrict__ dr,
           const int *__restrict__ vp, const int *__restrict__ dk, const double *__restrict__ Sx,
           const double *__restrict__ D, int ld, int dense_trans, int assoc,
           const double *__restrict__ d, double *__restrict__ C, long long sL, long long sC)
{
    const int L = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (L >= NL) return;
    const int e0 = lp[L], e1 = lp[L + 1];
    for (int c0 = 0; c0 < NC; c0 += 64)
    {
        const int c = c0 + lane;
        double acc = 0.0;
        for (int e = e0; e < e1; e++)
        {
            const int r = dr[e];
            const double sv = Sx[vp[e]];
            const int kd = dk ? dk[e] : -1;
            double dense = 0.0;
            if (c < NC) dense = dense_trans ? D[(long long) c * ld + r] : D[(long long) r * ld + c];
            double t;
            if (assoc == 0)
            {
                double coef = kd >= 0 ? d[kd] * sv : sv;
                t = coef * dense;
            }
            else
            {
                double dv = kd >= 0 ? d[kd] * dense : dense;
                t = sv * dv;
            }
            acc = acc + t;
        }
        if (c < NC) C[(long long) L * sL + (long long) c * sC] = acc;
    }
}

constexpr int GT = 64;   // output tile
constexpr int SK = 16;   // reduction slice
constexpr int PAD = 2;

__global__ void __launch_bounds__(256)
    k_ggemm(const GJob *__restrict__ jobs, const GTask *__restrict__ tasks,
            const int *__restrict__ pool, const double *__restrict__ buf0,
            const double *__restrict__ buf1, const double *__restrict__ d,
            double *__restrict__ C)
{
    __shared__ double Ps[SK][GT + PAD];
    __shared__ double Qs[SK][GT + PAD];
    const GJob job = jobs[blockIdx.x];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    for (int t = job.t0; t < job.t1; t++)
    {
        const GTask T = tasks[t];
        const double *P = (T.p_buf ? buf1 : buf0) + T.p_off;
        const double *Q = (T.q_buf ? buf1 : buf0) + T.q_off;
        for (int s0 = 0; s0 < T.S; s0 += SK)
        {
            for (int e = tid; e < SK * GT; e += 256)
            {
                const int ss = e / GT, ii = e % GT;
                const int s = s0 + ss;
                const int i = job.i0 + ii;
                double pv = 0.0, qv = 0.0;
                if (s < T.S)
                {
                    if (i < job.M)
                    {
                        const int mi = T.pmap < 0 ? i : pool[T.pmap + i];
                        if (mi >= 0)
                        {
                            const long long r = pool[T.prow + s];
                            pv = T.p_mode == 0 ? P[r * T.p_ld + mi] : P[(long long) mi * T.p_ld + r];
                        }
                    }
                    const int j = job.j0 + ii;
                    if (j < job.N)
                    {
                        const int mj = T.qmap < 0 ? j : pool[T.qmap + j];
                        if (mj >= 0)
                        {
                            const long long r = pool[T.qrow + s];
                            qv = T.q_mode == 0 ? Q[r * T.q_ld + mj] : Q[(long long) mj * T.q_ld + r];
                            if (d != nullptr && T.widx >= 0) qv = d[pool[T.widx + s]] * qv;
                        }
                    }
                }
                Ps[ss][ii] = pv;
                Qs[ss][ii] = qv;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < SK; kk += 4)
            {
                double af[2], bf[2];
#pragma unroll
                for (int a = 0; a < 2; a++) af[a] = Ps[kk + (lane >> 4)][wr * 32 + a * 16 + (lane & 15)];
#pragma unroll
                for (int b = 0; b < 2; b++) bf[b] = Qs[kk + (lane >> 4)][wc * 32 + b * 16 + (lane & 15)];
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
            }
            __syncthreads();
        }
    }
    double *Cb = C + job.c_off;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int i = job.i0 + wr * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int j = job.j0 + wc * 32 + b * 16 + (lane & 15);
                if (i < job.M && j < job.N) Cb[(long long) i * job.ldc + j] = acc[a][b][v];
            }
}

} // namespace

void pull_build(PullList &P, const std::vector<int> &seg, const std::vector<int> &idx)
{
    P.n = (int) seg.size() - 1;
    P.seg = upload(seg.data(), (int64_t) seg.size());
    P.idx = upload(idx.data(), (int64_t) idx.size());
}

void pull_sum(const PullList &P, const double *src, double *out, hipStream_t s)
{
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_pull_sum, dim3(grid_for(P.n)), dim3(256), 0, s, P.n, P.seg.p, P.idx.p, src,
                       out);
    SDL_HIP(hipGetLastError());
}

void sdmm_build(SdmmPlan &P, const std::vector<int> &lp, const std::vector<int> &dr,
                const std::vector<int> &vp, const std::vector<int> &dk)
{
    P.NL = (int) lp.size() - 1;
    P.nent = (long long) dr.size();
    P.lp = upload(lp.data(), (int64_t) lp.size());
    P.dr = upload(dr.data(), (int64_t) dr.size());
    P.vp = upload(vp.data(), (int64_t) vp.size());
    P.has_dk = !dk.empty();
    if (P.has_dk) P.dk = upload(dk.data(), (int64_t) dk.size());
}

void sdmm_run(const SdmmPlan &P, const double *Sx, const double *D, const double *d, double *C,
              hipStream_t s)
{
    if (P.NL <= 0 || P.NC <= 0) return;
    hipLaunchKernelGGL(k_sdmm, dim3((unsigned) ((P.NL + 3) / 4)), dim3(256), 0, s, P.NL, P.NC,
                       P.lp.p, P.dr.p, P.vp.p, (d && P.has_dk) ? P.dk.p : (const int *) nullptr, Sx,
                       D, P.ld, (int) P.dense_trans, P.assoc, d, C, P.sL, P.sC);
    SDL_HIP(hipGetLastError());
}

int GemmBuilder::add_ints(const std::vector<int> &v)
{
    int off = (int) pool.size();
    pool.insert(pool.end(), v.begin(), v.end());
    return off;
}

void GemmBuilder::add_output(long long c_off, int ldc, int M, int N, int t0, int t1)
{
    if (M <= 0 || N <= 0) return;
    for (int t = t0; t < t1; t++) flops += 2ll * tasks[t].S * M * N;
    for (int i0 = 0; i0 < M; i0 += GT)
        for (int j0 = 0; j0 < N; j0 += GT)
        {
            GJob j;
            j.c_off = c_off;
            j.ldc = ldc;
            j.M = M;
            j.N = N;
            j.i0 = i0;
            j.j0 = j0;
            j.t0 = t0;
            j.t1 = t1;
            jobs.push_back(j);
        }
}

void gemm_upload(GemmPlan &P, const GemmBuilder &B)
{
    P.njobs = (int) B.jobs.size();
    P.flops = B.flops;
    P.tasks = upload(B.tasks.data(), (int64_t) B.tasks.size());
    P.jobs = upload(B.jobs.data(), (int64_t) B.jobs.size());
    P.pool = upload(B.pool.data(), (int64_t) B.pool.size());
    sync();
}

void gemm_run(const GemmPlan &P, const double *buf0, const double *buf1, const double *d,
              double *C, hipStream_t s)
{
    if (P.njobs <= 0) return;
    hipLaunchKernelGGL(k_ggemm, dim3((unsigned) P.njobs), dim3(256), 0, s, P.jobs.p, P.tasks.p,
                       P.pool.p, buf0, buf1, d, C);
    SDL_HIP(hipGetLastError());
}

} // namespace sdl
