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

// pos null: the listed outputs are all outputs, in order.  Sources go four at
// a time with unconditional loads (indices past the output's last source
// repeat its first, whose value is then dropped), so four gathers are in
// flight per thread.
__global__ __launch_bounds__(256) void k_pull_sum(int n, const int *__restrict__ pos,
                                                 const int *__restrict__ seg,
                                                 const int *__restrict__ idx,
                                                 const double *__restrict__ src,
                                                 double *__restrict__ out)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s1 = seg[e + 1];
    double acc = 0.0;
    for (int s = seg[e]; s < s1; s += 4)
    {
        const int cnt = s1 - s;
        int ix[4];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) ix[u] = idx[u < cnt ? s + u : s];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = src[ix[u]];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (u < cnt) acc = acc + v[u];
    }
    out[pos ? pos[e] : e] = acc;
}

// One workgroup per chunk of single-source outputs: out = 0 + src (the sum of
// one term from +0, so a -0 source gives +0 as in the ordered sum).
__global__ void __launch_bounds__(256)
    k_pull_copy(const int *__restrict__ co, const int *__restrict__ cs, const int *__restrict__ cl,
                const double *__restrict__ src, double *__restrict__ out)
{
    const int c = blockIdx.x;
    const int o = co[c], sidx = cs[c], len = cl[c];
    for (int i = threadIdx.x; i < len; i += 256) out[o + i] = 0.0 + src[sidx + i];
}

// Exclusive prefix of v over the G threads of a workgroup: wave shuffles,
// then the G / 64 wave totals through LDS (wsum[G / 64]).
template <int G> __device__ __forceinline__ int block_excl_scan(int v, int *wsum)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; w++) base += wsum[w];
    return base + x - v;
}

// Row coalesce (RowCoalesce in pd_kernels.h): one workgroup of G threads per
// output row.  LDS: the dense row (ncolw doubles), the touched-column mask
// (ncolw / 32 words), each mask word's output rank, the scan's wave totals.
// NB sources are loaded at a time, H cells per thread per source, all loads
// unconditional (clamped to the source's last cell) so they are in flight
// together; the sources are then added in order, one barrier apart.
template <int G, int NB, int H>
__global__ __launch_bounds__(G) void k_coalesce_rows(
    int ncolw, const long long *__restrict__ obase, const int *__restrict__ sp,
    const long long *__restrict__ sstart, const int *__restrict__ slen,
    const int *__restrict__ scp, const int *__restrict__ cpool, const double *__restrict__ raw,
    double *__restrict__ out)
{
    extern __shared__ double rowc_lds[];
    double *dense = rowc_lds;
    const int nw = ncolw >> 5;
    unsigned *mask = (unsigned *) (dense + ncolw);
    int *wpre = (int *) (mask + nw);
    int *wsum = wpre + nw;
    const int t = threadIdx.x;
    const int row = blockIdx.x;
    for (int c = t; c < ncolw; c += G) dense[c] = 0.0;
    for (int w = t; w < nw; w += G) mask[w] = 0u;
    const int s0 = sp[row], s1 = sp[row + 1];
    __syncthreads();
    for (int sb = s0; sb < s1; sb += NB)
    {
        double v[NB][H];
        int col[NB][H];
        int len[NB];
#pragma unroll
        for (int u = 0; u < NB; u++)
        {
            const int s = min(sb + u, s1 - 1);
            const long long base = sstart[s];
            const int cpo = scp[s];
            len[u] = slen[s];
#pragma unroll
            for (int h = 0; h < H; h++)
            {
                const int j = min(t + G * h, len[u] - 1);
                v[u][h] = raw[base + j];
                col[u][h] = cpool[cpo + j];
            }
        }
#pragma unroll
        for (int u = 0; u < NB; u++)
        {
            if (sb + u < s1)  // the same for every thread
            {
#pragma unroll
                for (int h = 0; h < H; h++)
                {
                    if (t + G * h < len[u])
                    {
                        const int c = col[u][h];
                        dense[c] = dense[c] + v[u][h];
                        atomicOr(&mask[c >> 5], 1u << (c & 31));
                    }
                }
                __syncthreads();
            }
        }
    }
    // output rank of every mask word (nw <= 256 <= G: one word per thread)
    const int pc = t < nw ? __popc(mask[t]) : 0;
    const int ex = block_excl_scan<G>(pc, wsum);
    if (t < nw) wpre[t] = ex;
    __syncthreads();
    const long long ob = obase[row];
    for (int c = t; c < ncolw; c += G)
    {
        const unsigned mw = mask[c >> 5];
        const unsigned bit = 1u << (c & 31);
        if (mw & bit) out[ob + wpre[c >> 5] + __popc(mw & (bit - 1u))] = dense[c];
    }
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

constexpr int GT = 64;      // output tile
constexpr int PAD = 2;
constexpr int XCD = 8;      // workgroups are dealt round-robin to the XCDs
constexpr int XCD_RUN = 4;  // consecutive tiles kept on one XCD (shared panels)

// One SK x 64 slice of a task's P and Q operands into registers: thread tid
// owns tile column ii = tid % 64 and slice rows sb, sb + 4, ... (sb = tid / 64),
// so each wavefront reads 64 consecutive columns of a row.  The slice's row
// indices come first (direct, or pool lookups issued together), then every
// value load, each at a clamped valid address with out-of-range values
// selected to 0 afterwards, so no load waits behind a branch.
template <int SK>
__device__ __forceinline__ void load_slice(const GJob &job, const GTask &T, int s0,
                                           const int *__restrict__ pool,
                                           const double *__restrict__ buf0,
                                           const double *__restrict__ buf1,
                                           const double *__restrict__ w, int tid,
                                           double pr[SK / 4], double qr[SK / 4],
                                           double wr[SK / 4])
{
    constexpr int U = SK / 4;
    const double *P = (T.p_buf ? buf1 : buf0) + T.p_off;
    const double *Q = (T.q_buf ? buf1 : buf0) + T.q_off;
    const int ii = tid & 63, sb = tid >> 6;
    const int i = job.i0 + ii, j = job.j0 + ii;
    int mi = -1, mj = -1;
    if (i < job.M) mi = T.pmap < 0 ? i : pool[T.pmap + i];
    if (j < job.N) mj = T.qmap < 0 ? j : pool[T.qmap + j];
    int rp[U], rq[U];
    if (T.direct & 1)
    {
#pragma unroll
        for (int u = 0; u < U; u++) rp[u] = T.p_r0 + min(s0 + sb + 4 * u, T.S - 1);
    }
    else
    {
#pragma unroll
        for (int u = 0; u < U; u++) rp[u] = pool[T.prow + min(s0 + sb + 4 * u, T.S - 1)];
    }
    if (T.direct & 2)
    {
#pragma unroll
        for (int u = 0; u < U; u++) rq[u] = T.q_r0 + min(s0 + sb + 4 * u, T.S - 1);
    }
    else
    {
#pragma unroll
        for (int u = 0; u < U; u++) rq[u] = pool[T.qrow + min(s0 + sb + 4 * u, T.S - 1)];
    }
    const long long mic = mi >= 0 ? mi : 0, mjc = mj >= 0 ? mj : 0;
    double pv[U], qv[U], wv[U];
#pragma unroll
    for (int u = 0; u < U; u++)
    {
        const long long r = rp[u], q = rq[u];
        pv[u] = P[T.p_mode == 0 ? r * T.p_ld + mic : mic * T.p_ld + r];
        qv[u] = Q[T.q_mode == 0 ? q * T.q_ld + mjc : mjc * T.q_ld + q];
    }
    if (w != nullptr && T.widx >= 0)
    {
#pragma unroll
        for (int u = 0; u < U; u++) wv[u] = w[T.widx + min(s0 + sb + 4 * u, T.S - 1)];
    }
    else
    {
#pragma unroll
        for (int u = 0; u < U; u++) wv[u] = 1.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
    {
        const bool in = s0 + sb + 4 * u < T.S;
        pr[u] = (in && mi >= 0) ? pv[u] : 0.0;
        qr[u] = (in && mj >= 0) ? qv[u] : 0.0;
        wr[u] = in ? wv[u] : 1.0;
    }
}

__global__ void k_gather_w(int n, const int *__restrict__ src, const double *__restrict__ d,
                           double *__restrict__ w)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) w[q] = d[src[q]];
}

// Grid position of job index j (jobs sorted by decreasing work): runs of
// XCD_RUN consecutive jobs go to one XCD, runs dealt round-robin.
__host__ __device__ inline int job_of_block(int bid)
{
    const int xcd = bid % XCD, slot = bid / XCD;
    return ((slot / XCD_RUN) * XCD + xcd) * XCD_RUN + slot % XCD_RUN;
}

inline int grid_for_jobs(int njobs)
{
    const int runs = (njobs + XCD_RUN - 1) / XCD_RUN;
    return (runs + XCD - 1) / XCD * XCD * XCD_RUN;
}

template <int SK>
__global__ void __launch_bounds__(256)
    k_ggemm(int njobs, const GJob *__restrict__ jobs, const GTask *__restrict__ tasks,
            const int *__restrict__ pool, const double *__restrict__ buf0,
            const double *__restrict__ buf1, const double *__restrict__ w,
            double *__restrict__ C)
{
    __shared__ double Ps[SK][GT + PAD];
    __shared__ double Qs[SK][GT + PAD];
    const int jb = job_of_block(blockIdx.x);
    if (jb >= njobs) return;
    const GJob job = jobs[jb];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wrow = wave >> 1, wcol = wave & 1;
    const int ii = tid & 63, sb = tid >> 6;
    // 16 x 16 sub-tiles of this wavefront that intersect the output block
    // (wave-uniform; the others are skipped)
    bool live[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
            live[a][b] = job.i0 + wrow * 32 + a * 16 < job.M && job.j0 + wcol * 32 + b * 16 < job.N;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    int t = job.t0;
    while (t < job.t1 && tasks[t].S <= 0) t++;
    int s0 = 0;
    double pr[SK / 4], qr[SK / 4], wr[SK / 4];
    // the current task stays in registers; the task table is read again only
    // when the job moves to its next task
    GTask T;
    if (t < job.t1)
    {
        T = tasks[t];
        load_slice<SK>(job, T, 0, pool, buf0, buf1, w, tid, pr, qr, wr);
    }
    while (t < job.t1)
    {
        // the d factor scales Q's row before the product: ((d Q) entries)
#pragma unroll
        for (int u = 0; u < SK / 4; u++)
        {
            Ps[sb + 4 * u][ii] = pr[u];
            Qs[sb + 4 * u][ii] = wr[u] * qr[u];
        }
        __syncthreads();
        int nt = t, ns0 = s0 + SK;
        if (ns0 >= T.S)
        {
            ns0 = 0;
            for (nt = t + 1; nt < job.t1 && tasks[nt].S <= 0; nt++) {}
            if (nt < job.t1) T = tasks[nt];
        }
        // the next slice's loads are in flight during this slice's MFMAs
        if (nt < job.t1) load_slice<SK>(job, T, ns0, pool, buf0, buf1, w, tid, pr, qr, wr);
#pragma unroll
        for (int kk = 0; kk < SK; kk += 4)
        {
            double af[2], bf[2];
#pragma unroll
            for (int a = 0; a < 2; a++) af[a] = Ps[kk + (lane >> 4)][wrow * 32 + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 2; b++) bf[b] = Qs[kk + (lane >> 4)][wcol * 32 + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    if (live[a][b])
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
        t = nt;
        s0 = ns0;
    }
    double *Cb = C + job.c_off;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int i = job.i0 + wrow * 32 + a * 16 + (lane >> 4) + 4 * v;
                const int j = job.j0 + wcol * 32 + b * 16 + (lane & 15);
                if (i < job.M && j < job.N) Cb[(long long) i * job.ldc + j] = acc[a][b][v];
            }
}

int g_gemm_slice = 16;  // reduction slice depth of k_ggemm (16 or 32)

} // namespace

void pull_build(PullList &P, const std::vector<int> &seg, const std::vector<int> &idx)
{
    constexpr int MIN_RUN = 16, CHUNK = 4096;
    const int n = (int) seg.size() - 1;
    P.n = n;
    std::vector<int> co, cs, cl, pos, rseg{0}, ridx;
    auto single = [&](int e) { return seg[e + 1] - seg[e] == 1; };
    auto rest = [&](int e) {
        pos.push_back(e);
        for (int s = seg[e]; s < seg[e + 1]; s++) ridx.push_back(idx[s]);
        rseg.push_back((int) ridx.size());
    };
    int e = 0;
    while (e < n)
    {
        int f = e;
        if (single(e))
            while (f + 1 < n && single(f + 1) && idx[seg[f + 1]] == idx[seg[f]] + 1) f++;
        const int len = f - e + 1;
        if (single(e) && len >= MIN_RUN)
        {
            for (int o = 0; o < len; o += CHUNK)
            {
                co.push_back(e + o);
                cs.push_back(idx[seg[e]] + o);
                cl.push_back(std::min(CHUNK, len - o));
            }
        }
        else
        {
            for (int g = e; g <= f; g++) rest(g);
        }
        e = f + 1;
    }
    P.nchunks = (int) co.size();
    if (P.nchunks)
    {
        P.co = upload(co.data(), (int64_t) co.size());
        P.cs = upload(cs.data(), (int64_t) cs.size());
        P.cl = upload(cl.data(), (int64_t) cl.size());
    }
    P.nrest = (int) pos.size();
    if (P.nrest)
    {
        // every output listed in order: no position array
        if (P.nrest < n) P.pos = upload(pos.data(), (int64_t) pos.size());
        P.seg = upload(rseg.data(), (int64_t) rseg.size());
        P.idx = upload(ridx.data(), (int64_t) std::max<size_t>(ridx.size(), 1));
    }
}

void pull_sum(const PullList &P, const double *src, double *out, hipStream_t s)
{
    if (P.n <= 0) return;
    if (P.nchunks)
        hipLaunchKernelGGL(k_pull_copy, dim3(P.nchunks), dim3(256), 0, s, P.co.p, P.cs.p, P.cl.p,
                           src, out);
    if (P.nrest)
        hipLaunchKernelGGL(k_pull_sum, dim3(grid_for(P.nrest)), dim3(256), 0, s, P.nrest, P.pos.p,
                           P.seg.p, P.idx.p, src, out);
    SDL_HIP(hipGetLastError());
}

int g_coalesce_mode = 0;

namespace {
template <int G, int NB, int H>
void rowc_launch(const RowCoalesce &P, const double *raw, double *out, size_t lds, hipStream_t s)
{
    static bool attr = false;
    auto fn = k_coalesce_rows<G, NB, H>;
    if (!attr)
    {
        SDL_HIP(hipFuncSetAttribute((const void *) fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int) (ROWC_MAX_N * 8 + ROWC_MAX_N / 4 + 64)));
        attr = true;
    }
    hipLaunchKernelGGL(fn, dim3((unsigned) P.nrows), dim3(G), lds, s, P.ncolw, P.obase.p,
                       P.sp.p, P.sstart.p, P.slen.p, P.scp.p, P.cpool.p, raw, out);
}
} // namespace

void rowc_run(const RowCoalesce &P, const double *raw, double *out, hipStream_t s)
{
    if (P.nrows <= 0) return;
    SDL_REQUIRE(P.ncolw > 0 && P.ncolw <= ROWC_MAX_N && P.ncolw % 32 == 0,
                "row coalesce: column count out of range");
    // dense row, mask, word ranks, up to 16 wave totals
    const size_t lds = (size_t) P.ncolw * 8 + (size_t) (P.ncolw / 32) * 8 + 64;
    // P.h: 256-cell chunks of the longest source row
    switch (P.h)
    {
    case 1:
    case 2: rowc_launch<512, 16, 1>(P, raw, out, lds, s); break;
    case 4: rowc_launch<512, 8, 2>(P, raw, out, lds, s); break;
    case 8: rowc_launch<512, 4, 4>(P, raw, out, lds, s); break;
    case 16: rowc_launch<512, 2, 8>(P, raw, out, lds, s); break;
    case 32: rowc_launch<512, 1, 16>(P, raw, out, lds, s); break;
    default: SDL_REQUIRE(false, "row coalesce: source rows too long");
    }
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

namespace {
// rows pool[off .. off + S) are r0, r0 + 1, ...
bool contiguous(const std::vector<int> &pool, int off, int S, int &r0)
{
    if (S <= 0) return false;
    r0 = pool[off];
    for (int s = 1; s < S; s++)
        if (pool[off + s] != r0 + s) return false;
    return true;
}
} // namespace

void gemm_upload(GemmPlan &P, const GemmBuilder &B)
{
    P.njobs = (int) B.jobs.size();
    P.flops = B.flops;
    std::vector<GTask> tasks = B.tasks;
    // column maps that are the identity over every output tile they feed
    std::vector<int> pm_len(tasks.size(), 0), qm_len(tasks.size(), 0);
    for (const GJob &j : B.jobs)
        for (int t = j.t0; t < j.t1; t++)
        {
            pm_len[t] = std::max(pm_len[t], j.M);
            qm_len[t] = std::max(qm_len[t], j.N);
        }
    auto identity = [&](int off, int len) {
        for (int i = 0; i < len; i++)
            if (B.pool[off + i] != i) return false;
        return true;
    };
    std::vector<int> wsrc;
    for (size_t t = 0; t < tasks.size(); t++)
    {
        GTask &T = tasks[t];
        if (T.pmap >= 0 && identity(T.pmap, pm_len[t])) T.pmap = -1;
        if (T.qmap >= 0 && identity(T.qmap, qm_len[t])) T.qmap = -1;
        T.direct = 0;
        if (T.pmap < 0 && contiguous(B.pool, T.prow, T.S, T.p_r0)) T.direct |= 1;
        if (T.qmap < 0 && contiguous(B.pool, T.qrow, T.S, T.q_r0)) T.direct |= 2;
        if (T.widx >= 0)
        {
            const int pos = (int) wsrc.size();
            for (int s = 0; s < T.S; s++) wsrc.push_back(B.pool[T.widx + s]);
            T.widx = pos;
        }
    }
    P.nw = (int) wsrc.size();
    if (P.nw > 0)
    {
        P.wsrc = upload(wsrc.data(), (int64_t) wsrc.size());
        P.w.alloc(wsrc.size());
    }
    P.tasks = upload(tasks.data(), (int64_t) tasks.size());
    // longest jobs first (the grid drains evenly); ties keep block order so
    // the tiles of one block stay adjacent
    std::vector<GJob> jobs = B.jobs;
    std::vector<long long> work(tasks.size() + 1, 0);
    for (size_t t = 0; t < tasks.size(); t++) work[t + 1] = work[t] + std::max(tasks[t].S, 0);
    std::stable_sort(jobs.begin(), jobs.end(), [&](const GJob &a, const GJob &b) {
        return work[a.t1] - work[a.t0] > work[b.t1] - work[b.t0];
    });
    P.jobs = upload(jobs.data(), (int64_t) jobs.size());
    P.pool = upload(B.pool.data(), (int64_t) B.pool.size());
    sync();
}

void gemm_run(const GemmPlan &P, const double *buf0, const double *buf1, const double *d,
              double *C, hipStream_t s)
{
    if (P.njobs <= 0) return;
    const double *w = nullptr;
    if (d != nullptr && P.nw > 0)
    {
        hipLaunchKernelGGL(k_gather_w, dim3(grid_for(P.nw)), dim3(256), 0, s, P.nw, P.wsrc.p, d,
                           P.w.p);
        SDL_HIP(hipGetLastError());
        w = P.w.p;
    }
    const dim3 grid((unsigned) grid_for_jobs(P.njobs));
    if (g_gemm_slice == 32)
        hipLaunchKernelGGL(k_ggemm<32>, grid, dim3(256), 0, s, P.njobs, P.jobs.p, P.tasks.p,
                           P.pool.p, buf0, buf1, w, C);
    else
        hipLaunchKernelGGL(k_ggemm<16>, grid, dim3(256), 0, s, P.njobs, P.jobs.p, P.tasks.p,
                           P.pool.p, buf0, buf1, w, C);
    SDL_HIP(hipGetLastError());
}

} // namespace sdl

extern "C" {
#pragma GCC visibility push(default)
// Reduction slice depth of the block-product GEMM (16 or 32); returns the
// previous value, or -1 for an unsupported depth.
int sdl_block_gemm_slice(int sk)
{
    if (sk != 16 && sk != 32) return -1;
    const int prev = sdl::g_gemm_slice;
    sdl::g_gemm_slice = sk;
    return prev;
}
// Coalesce of block-product partials chosen at alloc: 0 = by output row
// where it applies, 1 = per-cell pull lists; returns the previous mode, or -1
// for an unknown one.
int sdl_block_coalesce_mode(int mode)
{
    if (mode != 0 && mode != 1) return -1;
    const int prev = sdl::g_coalesce_mode;
    sdl::g_coalesce_mode = mode;
    return prev;
}
#pragma GCC visibility pop
}
