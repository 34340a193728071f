// Gustavson SpGEMM on CDNA4: C = kron(I_blocks, L) . diag(d) . A.
//
// Symbolic (alloc): per-row term bound -> rows binned by size class on the
// device -> one workgroup per row with an LDS hash table (64..8192 slots), an
// LDS bitmap (dense rows; columns come out sorted for free) or a global bitmap
// (huge rows) -> count pass -> exclusive scan -> fill pass with the same
// accumulators sized from the exact count -> bitonic sort of each hash table.
//
// Refill schedule: the reference's plan policy (terms <= 3 nnz(C)) picks a
// pull schedule, where every output entry owns the list of its terms in
// ascending k; otherwise a row accumulator walks each row's k in order.
// Either way every entry is the left fold
//     C[i,j] = (((0 + (L[i,k1] d[k1]) A[k1,j]) + (L[i,k2] d[k2]) A[k2,j]) + ...)
// in ascending k with individually rounded products and sums (the library is
// built with -ffp-contract=off), which is gustavson.c:127-135 / :220-228 bit
// for bit.
#include "spgemm.h"

#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <climits>
#include <vector>

namespace sdl {

int g_sym_window = 1;
int g_sym_window_lanes = 0;
int g_rowacc_window = 1;
int g_sym_groups = 1;

namespace {

// 0 empty, 1..8 hash 64..8192, 9 LDS bitmap, 10 global bitmap, 11..16 span
// windows of 256..8192 columns
constexpr int NCLASS = 17;
constexpr int CLS_BITMAP = 9;
constexpr int CLS_GLOBAL = 10;
constexpr int CLS_WIN0 = 11;
constexpr int WIN_MIN = 256, WIN_MAX = 8192;
constexpr int HASH_MAX = 8192;

int g_lds_bitmap_words = -1;    // largest bitmap held in LDS by one workgroup

int lds_bitmap_words()
{
    if (g_lds_bitmap_words < 0)
    {
        int dev = 0;
        SDL_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        SDL_HIP(hipGetDeviceProperties(&prop, dev));
        size_t lds = prop.sharedMemPerBlock;
        if (lds > 160 * 1024) lds = 160 * 1024;
        // keep 8 KB for the kernel's static arrays
        g_lds_bitmap_words = (int) ((lds - 8192) / 4);
    }
    return g_lds_bitmap_words;
}

// ------------------------------------------------------------- helpers ----
template <int G> __device__ __forceinline__ int block_excl_scan_int(int x, int *total, int *sh_w)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        int y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (G == 64)
    {
        *total = __shfl(inc, 63, 64);
        return inc - x;
    }
    if (lane == 63) sh_w[w] = inc;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < G / 64; i++)
    {
        int s = sh_w[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - x;
}

template <int G> __device__ __forceinline__ long long block_sum_ll(long long x, long long *sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (G == 64) return x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = x;
    __syncthreads();
    long long t = 0;
#pragma unroll
    for (int i = 0; i < G / 64; i++) t += sh[i];
    __syncthreads();
    return t;
}

// Shared state of the load-balanced term expansion of one output row: a chunk
// of up to G entries of L's row with their A-row start and exclusive offsets.
template <int G> struct Expand
{
    int start[G];
    int k[G];
    int q[G];
    int off[G + 1];
    int wsum[G / 64];
};

// Calls f(t, q, k) for every term of output row (b, r) -- t indexes A's
// entries, q indexes L's entries, k is the A row.  All G threads must call it.
template <int G, typename F>
__device__ __forceinline__ void for_each_term(Expand<G> &ex, int q0, int q1, int koff,
                                              const int *__restrict__ Li,
                                              const int *__restrict__ Ap, F &&f)
{
    for (int qb = q0; qb < q1; qb += G)
    {
        int q = qb + (int) threadIdx.x;
        int len = 0, st = 0, k = 0;
        if (q < q1)
        {
            k = Li[q] + koff;
            st = Ap[k];
            len = Ap[k + 1] - st;
        }
        int tot;
        int ex_off = block_excl_scan_int<G>(len, &tot, ex.wsum);
        ex.start[threadIdx.x] = st;
        ex.k[threadIdx.x] = k;
        ex.q[threadIdx.x] = q;
        ex.off[threadIdx.x] = ex_off;
        if (threadIdx.x == 0) ex.off[G] = tot;
        __syncthreads();
        for (int fi = (int) threadIdx.x; fi < tot; fi += G)
        {
            // last slot s with off[s] <= fi (slots with len 0 share offsets)
            int lo = 0, hi = G - 1;
            while (lo < hi)
            {
                int mid = (lo + hi + 1) >> 1;
                if (ex.off[mid] <= fi) lo = mid;
                else hi = mid - 1;
            }
            f(ex.start[lo] + (fi - ex.off[lo]), ex.q[lo], ex.k[lo]);
        }
        __syncthreads();
    }
}

// The chunk setup of for_each_term as a separate step (for kernels that
// process several terms per thread at once): L entries qb .. qb+G-1 of the
// row with their A-row starts and exclusive term offsets; returns the chunk's
// term count.  All G threads must call it; it ends with a barrier.
template <int G>
__device__ __forceinline__ int expand_chunk(Expand<G> &ex, int qb, int q1, int koff,
                                            const int *__restrict__ Li,
                                            const int *__restrict__ Ap)
{
    int q = qb + (int) threadIdx.x;
    int len = 0, st = 0, k = 0;
    if (q < q1)
    {
        k = Li[q] + koff;
        st = Ap[k];
        len = Ap[k + 1] - st;
    }
    int tot;
    int ex_off = block_excl_scan_int<G>(len, &tot, ex.wsum);
    ex.start[threadIdx.x] = st;
    ex.k[threadIdx.x] = k;
    ex.q[threadIdx.x] = q;
    ex.off[threadIdx.x] = ex_off;
    if (threadIdx.x == 0) ex.off[G] = tot;
    __syncthreads();
    return tot;
}

// Term fi of the current chunk (nq = its number of L entries): A entry t,
// L entry q, A row k.
template <int G>
__device__ __forceinline__ void expand_term(const Expand<G> &ex, int nq, int fi, int &t, int &q,
                                            int &k)
{
    int lo = 0, hi = nq - 1;
    while (lo < hi)
    {
        int mid = (lo + hi + 1) >> 1;
        if (ex.off[mid] <= fi) lo = mid;
        else hi = mid - 1;
    }
    t = ex.start[lo] + (fi - ex.off[lo]);
    q = ex.q[lo];
    k = ex.k[lo];
}

// Calls f(col) for the column of every term of output row (b, r), U terms
// per thread at a time with their column loads issued together (slots past
// a chunk repeat its last term and are dropped, so no load sits behind a
// branch).  All G threads must call it.
template <int G, int U, typename F>
__device__ __forceinline__ void for_each_col(Expand<G> &ex, int q0, int q1, int koff,
                                             const int *__restrict__ Li,
                                             const int *__restrict__ Ap,
                                             const int *__restrict__ Ai, F &&f)
{
    for (int qb = q0; qb < q1; qb += G)
    {
        const int tot = expand_chunk<G>(ex, qb, q1, koff, Li, Ap);
        const int nq = min(G, q1 - qb);
        for (int f0 = (int) threadIdx.x; f0 < tot; f0 += U * G)
        {
            int col[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int fi = f0 + u * G;
                int t, q, k;
                expand_term<G>(ex, nq, fi < tot ? fi : tot - 1, t, q, k);
                const int a = Ai[t];
                col[u] = fi < tot ? a : -1;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (col[u] >= 0) f(col[u]);
        }
        __syncthreads();
    }
}

// Calls f(col) for the column of every term of output row (b, r): groups of
// sg lanes (sg divides G) take one L entry at a time and stride over its A
// row, U loads in flight per lane (a slot past the row reloads its last
// entry and is dropped).  No workgroup barrier; the terms of one L entry are
// visited by one group.
template <int G, int U = 4, typename F>
__device__ __forceinline__ void for_each_col_groups(int sg, int q0, int q1, int koff,
                                                    const int *__restrict__ Li,
                                                    const int *__restrict__ Ap,
                                                    const int *__restrict__ Ai, F &&f)
{
    const int grp = (int) threadIdx.x / sg, ngrp = G / sg, sl = (int) threadIdx.x - grp * sg;
    for (int q = q0 + grp; q < q1; q += ngrp)
    {
        const int k = Li[q] + koff;
        const int st = Ap[k], en = Ap[k + 1];
        for (int t = st + sl; t < en; t += U * sg)
        {
            int col[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int tu = t + u * sg;
                const int a = Ai[min(tu, en - 1)];
                col[u] = tu < en ? a : -1;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (col[u] >= 0) f(col[u]);
        }
    }
}

template <int S> __device__ __forceinline__ unsigned hash_slot(int col)
{
    constexpr int LOG2 = __builtin_ctz(S);
    return ((unsigned) col * 2654435761u) >> (32 - LOG2);
}

// Insert col; returns 1 when col took an empty slot, 0 when it was there
// already, -1 when the probe sequence ran over the whole table.
template <int S> __device__ __forceinline__ int hash_insert(int *tab, int col)
{
    unsigned h = hash_slot<S>(col);
    for (int probe = 0; probe < S; probe++)
    {
        int prev = atomicCAS(&tab[h], -1, col);
        if (prev == -1) return 1;
        if (prev == col) return 0;
        h = (h + 1) & (S - 1);
    }
    return -1;
}

// Ascending bitonic sort of S unsigned keys in LDS by G threads.
template <int G, int S> __device__ __forceinline__ void bitonic_sort(unsigned *a)
{
    for (int k = 2; k <= S; k <<= 1)
    {
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int p = (int) threadIdx.x; p < S / 2; p += G)
            {
                int i = (p / j) * 2 * j + (p % j);
                int ixj = i + j;
                unsigned x = a[i], y = a[ixj];
                bool up = (i & k) == 0;
                if ((x > y) == up)
                {
                    a[i] = y;
                    a[ixj] = x;
                }
            }
            __syncthreads();
        }
    }
}

// The same sort over the first n keys (n a power of two, at most the array).
template <int G> __device__ __forceinline__ void bitonic_sort_n(unsigned *a, int n)
{
    for (int k = 2; k <= n; k <<= 1)
    {
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int p = (int) threadIdx.x; p < n / 2; p += G)
            {
                int i = (p / j) * 2 * j + (p % j);
                int ixj = i + j;
                unsigned x = a[i], y = a[ixj];
                bool up = (i & k) == 0;
                if ((x > y) == up)
                {
                    a[i] = y;
                    a[ixj] = x;
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ int lower_bound_int(const int *a, int n, int x)
{
    int lo = 0, hi = n;
    while (lo < hi)
    {
        int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

struct RowCtx
{
    int q0, q1, koff;
};

__device__ __forceinline__ RowCtx row_ctx(int row, int Lm, const int *Lp, int nL)
{
    int b = row / Lm, r = row - b * Lm;
    RowCtx c;
    c.q0 = Lp[r];
    c.q1 = Lp[r + 1];
    c.koff = b * nL;
    return c;
}

// ------------------------------------------------------------ kernels ----
// Smallest and largest column of every row of A (one wavefront per row;
// rows need not be sorted).  Empty rows get (INT_MAX, -1).
__global__ void __launch_bounds__(256)
    k_arow_range(int Am, const int *__restrict__ Ap, const int *__restrict__ Ai,
                 int2 *__restrict__ range)
{
    const int64_t k = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= Am) return;
    int lo = INT_MAX, hi = -1;
    for (int t = Ap[k] + lane; t < Ap[k + 1]; t += 64)
    {
        const int col = Ai[t];
        lo = min(lo, col);
        hi = max(hi, col);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
    }
    if (lane == 0) range[k] = make_int2(lo, hi);
}

// Per output row: its term count, and (bound non-null) its nnz bound
// min(terms, column span): a row's columns all lie between the smallest and
// the largest column of its A rows (arange, from k_arow_range), so the span
// bounds the distinct count exactly.  With win_lo non-null also the span
// itself: its first column and width (0 for a row without terms).  16 lanes
// per row, four rows per wavefront.
__global__ void __launch_bounds__(256)
    k_row_terms(int64_t m_out, int Lm, const int *__restrict__ Lp, const int *__restrict__ Li,
                int nL, const int *__restrict__ Ap, const int2 *__restrict__ arange,
                int64_t *__restrict__ terms, int64_t *__restrict__ bound,
                int *__restrict__ win_lo = nullptr, int *__restrict__ win_span = nullptr)
{
    const int64_t row = (int64_t) blockIdx.x * 16 + (threadIdx.x >> 4);
    const int sub = threadIdx.x & 15;
    const bool live = row < m_out;
    int64_t t = 0;
    int cmin = INT_MAX, cmax = -1;
    if (live)
    {
        RowCtx c = row_ctx((int) row, Lm, Lp, nL);
        for (int q = c.q0 + sub; q < c.q1; q += 16)
        {
            const int k = Li[q] + c.koff;
            t += Ap[k + 1] - Ap[k];
            if (bound)
            {
                const int2 r = arange[k];
                cmin = min(cmin, r.x);
                cmax = max(cmax, r.y);
            }
        }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1)
    {
        t += __shfl_xor(t, off, 16);
        cmin = min(cmin, __shfl_xor(cmin, off, 16));
        cmax = max(cmax, __shfl_xor(cmax, off, 16));
    }
    if (!live || sub != 0) return;
    terms[row] = t;
    if (bound) bound[row] = t > 0 ? min(t, (int64_t) cmax - cmin + 1) : 0;
    if (win_lo)
    {
        win_lo[row] = t > 0 ? cmin : 0;
        win_span[row] = t > 0 ? cmax - cmin + 1 : 0;
    }
}

// Accumulator class of a row whose distinct count is at most cap: a hash
// table of the smallest power of two >= 2 cap slots (64 .. hash_max), or the
// LDS column bitmap when the row is dense enough (cap * dens >= nwords) or its
// table would pass hash_max, or a bitmap in global memory when the bitmap
// does not fit LDS either.  hash_max and dens are tuning parameters (SymTune).
__device__ __forceinline__ int classify(int64_t cap, int nwords, int lds_words,
                                        int hash_max = HASH_MAX, int dens = 8)
{
    if (cap <= 0) return 0;
    bool bm_ok = nwords <= lds_words;
    if (bm_ok && (cap * dens >= nwords || 2 * cap > hash_max)) return CLS_BITMAP;
    if (2 * cap <= hash_max)
    {
        int64_t s = 64;
        int c = 1;
        while (s < 2 * cap)
        {
            s <<= 1;
            c++;
        }
        return c;
    }
    return CLS_GLOBAL;
}

// Class of a row of `terms` terms whose columns lie in a span of `span`
// columns: the span window (CLS_WIN0 + j, W = 256 << j columns) when the span
// fits WIN_MAX and the window's slots, which it clears and scans, are at most
// twice the hash table the row would get or half its terms (a hash insert
// costs a compare-and-swap and probes, a window mark one store; the fill pass
// also sorts the table); against the column bitmap, when the window's slots
// exceed the bitmap's words by less than the row's terms (consecutive
// columns put up to 32 lanes on one bitmap word, whose atomic ORs
// serialise, where the window spreads them over the LDS banks).  Otherwise
// classify's choice.
__device__ __forceinline__ int classify_span(int64_t cap, int64_t span, int64_t terms, int nwords,
                                             int lds_words, int hash_max, int dens)
{
    if (cap <= 0) return 0;
    if (span > 0 && span <= WIN_MAX)
    {
        int j = 0;
        while ((WIN_MIN << j) < span) j++;
        const int64_t wcols = (int64_t) WIN_MIN << j;
        const bool bitmap =
            nwords <= lds_words && (cap * dens >= nwords || 2 * cap > hash_max);
        if (bitmap)
        {
            if (wcols <= (int64_t) nwords + terms) return CLS_WIN0 + j;
        }
        else
        {
            int64_t hs = 64;
            while (hs < 2 * cap) hs <<= 1;
            if (wcols <= 2 * hs || 2 * wcols <= terms) return CLS_WIN0 + j;
        }
    }
    return classify(cap, nwords, lds_words, hash_max, dens);
}

// mode 0: symbolic accumulator classes (classify with hash_max / dens, or
// classify_span where span is given); mode 1: 0 empty, 1 <= t1, 2 <= t2, 3
// larger
template <typename T>
__global__ void k_classify(int64_t m, const T *__restrict__ cap, int mode, int nwords,
                           int lds_words, int hash_max, int dens, int t1, int t2,
                           const int *__restrict__ span, const int64_t *__restrict__ terms,
                           unsigned char *__restrict__ cls, int *__restrict__ counts)
{
    __shared__ int hist[NCLASS];
    if (threadIdx.x < NCLASS) hist[threadIdx.x] = 0;
    __syncthreads();
    int64_t row = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (row < m)
    {
        int64_t v = (int64_t) cap[row];
        int c = mode == 0 ? (span ? classify_span(v, span[row], terms[row], nwords, lds_words,
                                                  hash_max, dens)
                                  : classify(v, nwords, lds_words, hash_max, dens))
                          : (v <= 0 ? 0 : (v <= t1 ? 1 : (v <= t2 ? 2 : 3)));
        cls[row] = (unsigned char) c;
        atomicAdd(&hist[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < NCLASS && hist[threadIdx.x]) atomicAdd(&counts[threadIdx.x], hist[threadIdx.x]);
}

// Row ids into their class's range: one atomic per (wavefront, class) --
// most rows of a product share a class, and per-row atomics on one counter
// serialise.
__global__ void k_scatter_rows(int64_t m, const unsigned char *__restrict__ cls,
                               int *__restrict__ cursor, int *__restrict__ rows,
                               const int *__restrict__ ids = nullptr)
{
    const int64_t row = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int c = row < m ? (int) cls[row] : 0;
    unsigned long long todo = __ballot(c != 0);
    while (todo)
    {
        const int leader = __ffsll((long long) todo) - 1;
        const int cl = __shfl(c, leader, 64);
        const unsigned long long same = __ballot(c == cl);
        int base = 0;
        if (lane == leader) base = atomicAdd(&cursor[cl], __popcll(same));
        base = __shfl(base, leader, 64);
        if (c == cl)
            rows[base + __popcll(same & ((1ull << lane) - 1ull))] = ids ? ids[row] : (int) row;
        todo &= ~same;
    }
}

// PASS 0: count distinct columns -> rowcnt.  PASS 1: write sorted columns.
template <int G, int S, int PASS>
__global__ void __launch_bounds__(G)
    k_sym_hash(const int *__restrict__ rows, int Lm, const int *__restrict__ Lp,
               const int *__restrict__ Li, int nL, const int *__restrict__ Ap,
               const int *__restrict__ Ai, int *__restrict__ rowcnt,
               const int64_t *__restrict__ Cp, int *__restrict__ Ci, int sg)
{
    __shared__ int tab[S];
    __shared__ Expand<G> ex;
    __shared__ long long red[G / 64];
    __shared__ int filled, ncomp;
    const int row = rows[blockIdx.x];
    for (int i = threadIdx.x; i < S; i += G) tab[i] = -1;
    if (threadIdx.x == 0)
    {
        filled = 0;
        ncomp = 0;
    }
    __syncthreads();
    RowCtx c = row_ctx(row, Lm, Lp, nL);
    // Tables hold at least twice their row's bound.  In the count pass a row
    // past S/2 distinct columns has exceeded its (predicted) bound and will be
    // counted again from its term bound, so its remaining terms are skipped:
    // without this, a row far above its table would probe the full table for
    // every new column.  Exact bounds never get there.
    // (lane groups per L entry when sg > 0, else the load-balanced expansion)
    auto put = [&](int col) {
        if (PASS == 0)
        {
            if (*(volatile int *) &filled > S / 2) return;
            if (hash_insert<S>(tab, col) == 1) atomicAdd(&filled, 1);
        }
        else
        {
            hash_insert<S>(tab, col);
        }
    };
    if (sg > 0) for_each_col_groups<G>(sg, c.q0, c.q1, c.koff, Li, Ap, Ai, put);
    else for_each_col<G, 4>(ex, c.q0, c.q1, c.koff, Li, Ap, Ai, put);
    __syncthreads();
    if (PASS == 0)
    {
        long long cnt = 0;
        for (int i = threadIdx.x; i < S; i += G) cnt += tab[i] != -1;
        cnt = block_sum_ll<G>(cnt, red);
        if (threadIdx.x == 0) rowcnt[row] = (int) cnt;
    }
    else
    {
        // the occupied slots (at most half the table) gathered to its front in
        // any order, padded with empty keys to a power of two n, and only
        // those n keys sorted (empty = 0xffffffff sorts last)
        static_assert(S % G == 0, "table slots per thread");
        constexpr int PER = S / G;
        const int lane = (int) threadIdx.x & 63;
        int v[PER];
#pragma unroll
        for (int i = 0; i < PER; i++) v[i] = tab[(int) threadIdx.x + i * G];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; i++)
        {
            const bool occ = v[i] != -1;
            const unsigned long long b = __ballot(occ);
            int at = 0;
            if (lane == 0 && b) at = atomicAdd(&ncomp, __popcll(b));
            at = __shfl(at, 0, 64);
            if (occ) tab[at + __popcll(b & ((1ull << lane) - 1ull))] = v[i];
        }
        __syncthreads();
        const int got = ncomp;
        int n = 1;
        while (n < got) n <<= 1;
        for (int i = got + (int) threadIdx.x; i < n; i += G) tab[i] = -1;
        __syncthreads();
        bitonic_sort_n<G>((unsigned *) tab, n);
        int64_t base = Cp[row];
        int cnt = (int) (Cp[row + 1] - base);
        cnt = min(cnt, got);
        for (int i = threadIdx.x; i < cnt; i += G) Ci[base + i] = tab[i];
    }
}

// Bitmap accumulator over all n columns, in LDS (GLOBAL = false) or in a
// per-workgroup slice of global scratch (GLOBAL = true, grid-strided rows).
template <int G, int PASS, bool GLOBAL>
__global__ void __launch_bounds__(G)
    k_sym_bitmap(const int *__restrict__ rows, int nrows, int nwords, int Lm,
                 const int *__restrict__ Lp, const int *__restrict__ Li, int nL,
                 const int *__restrict__ Ap, const int *__restrict__ Ai,
                 int *__restrict__ rowcnt, const int64_t *__restrict__ Cp,
                 int *__restrict__ Ci, unsigned *__restrict__ gscratch, int sg)
{
    extern __shared__ unsigned lds_bm[];
    __shared__ Expand<G> ex;
    __shared__ long long red[G / 64];
    __shared__ int wsum[G / 64];
    // In global scratch the bits are set by atomics (performed in L2), so the
    // words are cleared and read with agent-scope atomic stores / loads and
    // the phases are separated by device-scope fences.
    unsigned *bm = GLOBAL ? gscratch + (size_t) blockIdx.x * nwords : lds_bm;
    auto word = [&](int w) -> unsigned {
        if (!GLOBAL) return bm[w];
        return __hip_atomic_load(bm + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    for (int ri = blockIdx.x; ri < nrows; ri += (GLOBAL ? gridDim.x : nrows))
    {
        const int row = rows[ri];
        for (int i = threadIdx.x; i < nwords; i += G)
        {
            if (GLOBAL) __hip_atomic_store(bm + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else bm[i] = 0u;
        }
        if (GLOBAL) __threadfence();
        __syncthreads();
        RowCtx c = row_ctx(row, Lm, Lp, nL);
        // LDS bitmaps: lane groups per L entry (sg > 0), 16 loads in flight
        // per lane (one workgroup per CU: the in-flight loads set the read
        // rate over the rows' long A rows; 4 / 8 measured slower), or the
        // load-balanced expansion; global ones: the expansion with one load in
        // flight (their L2 atomics measured slower batched)
        auto put = [&](int col) { atomicOr(&bm[col >> 5], 1u << (col & 31)); };
        if (!GLOBAL && sg > 0) for_each_col_groups<G, 16>(sg, c.q0, c.q1, c.koff, Li, Ap, Ai, put);
        else for_each_col<G, GLOBAL ? 1 : 4>(ex, c.q0, c.q1, c.koff, Li, Ap, Ai, put);
        if (GLOBAL) __threadfence();
        __syncthreads();
        if (PASS == 0)
        {
            // lanes read consecutive words (no LDS bank conflicts)
            int cnt = 0;
            for (int w = (int) threadIdx.x; w < nwords; w += G) cnt += __popc(word(w));
            long long tot = block_sum_ll<G>((long long) cnt, red);
            if (threadIdx.x == 0) rowcnt[row] = (int) tot;
        }
        else
        {
            // wavefront v owns words [v * pw, (v + 1) * pw), 256 at a time,
            // four consecutive words per lane: the wavefronts' counts give
            // their output offsets, then each chunk's columns go out in
            // ascending order at the chunk's running offset plus the lane's
            // exclusive prefix of the chunk's counts (one prefix per 256 words)
            constexpr int R = 4;
            const int lane = (int) threadIdx.x & 63, wv = (int) threadIdx.x >> 6;
            const int pw = (nwords + G - 1) / G * 64;
            const int w0 = min(nwords, wv * pw), w1 = min(nwords, w0 + pw);
            int cnt = 0;
            for (int w = w0 + lane; w < w1; w += 64) cnt += __popc(word(w));
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
            if (lane == 0) wsum[wv] = cnt;
            __syncthreads();
            int64_t base = Cp[row];
            for (int v = 0; v < wv; v++) base += wsum[v];
            for (int wb = w0; wb < w1; wb += 64 * R)
            {
                const int wl = wb + lane * R;
                unsigned bits[R];
                int c = 0;
#pragma unroll
                for (int r = 0; r < R; r++)
                {
                    bits[r] = wl + r < w1 ? word(wl + r) : 0u;
                    c += __popc(bits[r]);
                }
                int inc = c;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1)
                {
                    const int y = __shfl_up(inc, off, 64);
                    if (lane >= off) inc += y;
                }
                int64_t at = base + inc - c;
#pragma unroll
                for (int r = 0; r < R; r++)
                {
                    unsigned b32 = bits[r];
                    while (b32)
                    {
                        const int b = __ffs(b32) - 1;
                        b32 &= b32 - 1;
                        Ci[at++] = (wl + r) * 32 + b;
                    }
                }
                base += __shfl(inc, 63, 64);
            }
        }
        __syncthreads();
        if (!GLOBAL) break;
    }
}

// Span-window accumulator of a row whose columns all lie in [lo, lo + W)
// (lo = its smallest column, from k_row_terms): an LDS mark per column of the
// window, set by plain stores -- any number of terms may mark one column, so
// there is no atomic, no probe and no overflow -- counted by wavefront ballots
// and, in the fill pass, read out in ascending order with no sort.  sg lanes
// take one L entry at a time and stride over its A row, four loads in flight
// (sg divides G).  A column outside the window cannot occur (the span bounds
// every column of the row) and would be dropped by the range check.
template <int G, int W, int PASS>
__global__ void __launch_bounds__(G)
    k_sym_window(const int *__restrict__ rows, int Lm, const int *__restrict__ Lp,
                 const int *__restrict__ Li, int nL, const int *__restrict__ Ap,
                 const int *__restrict__ Ai, const int *__restrict__ win_lo, int sg,
                 int *__restrict__ rowcnt, const int64_t *__restrict__ Cp,
                 int *__restrict__ Ci)
{
    static_assert(W % G == 0 && W / (G / 64) >= 64, "window slots per wavefront");
    constexpr int NW = G / 64, PER = W / NW;
    __shared__ int mark[W];
    __shared__ int wcnt[NW];
    const int wid = (int) threadIdx.x >> 6, lane = (int) threadIdx.x & 63;
    const int row = rows[blockIdx.x];
    const int lo = win_lo[row];
    for (int i = (int) threadIdx.x; i < W; i += G) mark[i] = 0;
    __syncthreads();
    const RowCtx c = row_ctx(row, Lm, Lp, nL);
    for_each_col_groups<G>(sg, c.q0, c.q1, c.koff, Li, Ap, Ai, [&](int col) {
        const unsigned o = (unsigned) (col - lo);
        if (o < (unsigned) W) mark[o] = 1;
    });
    __syncthreads();
    // wavefront w owns window slots [w * PER, (w + 1) * PER)
    const int s0 = wid * PER;
    int cnt = 0;
    for (int i = s0 + lane; i < s0 + PER; i += 64) cnt += __popcll(__ballot(mark[i] != 0));
    if (lane == 0) wcnt[wid] = cnt;
    __syncthreads();
    if (PASS == 0)
    {
        if (threadIdx.x == 0)
        {
            int tot = 0;
            for (int w = 0; w < NW; w++) tot += wcnt[w];
            rowcnt[row] = tot;
        }
    }
    else
    {
        // (the row's count came from the count pass; writes stay below the
        // row's end whatever the marks)
        int64_t pos = Cp[row];
        const int64_t end = Cp[row + 1];
        for (int w = 0; w < wid; w++) pos += wcnt[w];
        const unsigned long long below = (1ull << lane) - 1ull;
        for (int i = s0 + lane; i < s0 + PER; i += 64)
        {
            const bool on = mark[i] != 0;
            const unsigned long long b = __ballot(on);
            const int64_t at = pos + __popcll(b & below);
            if (on && at < end) Ci[at] = lo + i;
            pos += __popcll(b);
        }
    }
}

// ---------------------------------------------------- learned binning ----
__device__ __forceinline__ float lg(double x) { return (float) log2(1.0 + x); }

// Structural features of every output row: one wavefront per row, the lanes
// striding over the row's L entries (reading A's row bounds and first / last
// columns), then wavefront reductions.  The sums are of integers below 2^53,
// so they are exact in any order.
__global__ void __launch_bounds__(256)
    k_row_features(int64_t nrows, const int *__restrict__ list, int Lm, const int *__restrict__ Lp,
                   const int *__restrict__ Li, int nL, const int *__restrict__ Ap,
                   const int *__restrict__ Ai, int An, float *__restrict__ F)
{
    // feature row i describes output row list[i] (list null: row i)
    const int64_t i = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= nrows) return;
    const int64_t row = list ? list[i] : i;
    RowCtx c = row_ctx((int) row, Lm, Lp, nL);
    const int deg = c.q1 - c.q0;
    double terms = 0, maxlen = 0, minlen = 1e300, sumsq = 0, sspan = 0;
    long long cmin = LLONG_MAX, cmax = -1, kmin = LLONG_MAX, kmax = -1;
    for (int q = c.q0 + lane; q < c.q1; q += 64)
    {
        const int k = Li[q] + c.koff;
        const int st = Ap[k], en = Ap[k + 1];
        const double len = en - st;
        kmin = min(kmin, (long long) k);
        kmax = max(kmax, (long long) k);
        if (len > 0)
        {
            terms += len;
            maxlen = fmax(maxlen, len);
            minlen = fmin(minlen, len);
            sumsq += len * len;
            const long long f = Ai[st], l = Ai[en - 1];
            cmin = min(cmin, f);
            cmax = max(cmax, l);
            sspan += (double) (l - f + 1);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        terms += __shfl_xor(terms, off, 64);
        sumsq += __shfl_xor(sumsq, off, 64);
        sspan += __shfl_xor(sspan, off, 64);
        maxlen = fmax(maxlen, __shfl_xor(maxlen, off, 64));
        minlen = fmin(minlen, __shfl_xor(minlen, off, 64));
        cmin = min(cmin, __shfl_xor(cmin, off, 64));
        cmax = max(cmax, __shfl_xor(cmax, off, 64));
        kmin = min(kmin, __shfl_xor(kmin, off, 64));
        kmax = max(kmax, __shfl_xor(kmax, off, 64));
    }
    if (lane != 0) return;
    if (terms == 0) minlen = 0;
    (void) An;  // widths are the row's column span, never n
    const double span = terms > 0 ? (double) (cmax - cmin + 1) : 0.0;
    const double kspan = deg > 0 ? (double) (kmax - kmin + 1) : 0.0;
    const double mean = deg > 0 ? terms / deg : 0.0;
    float *f = F + i * 16;
    f[0] = lg(terms);
    f[1] = lg(deg);
    f[2] = lg(maxlen);
    f[3] = lg(minlen);
    f[4] = lg(mean);
    f[5] = lg(span);
    f[6] = lg(span > 0 ? span * (1.0 - exp(-terms / span)) : 0.0);
    f[7] = lg(sspan);
    f[8] = lg(fmin(terms, sspan));
    f[9] = lg(fmin(terms, span));
    f[10] = lg(kspan);
    f[11] = kspan > 0 ? (float) (deg / kspan) : 0.0f;
    f[12] = lg(sqrt(sumsq));
    f[13] = lg(deg > 0 ? sspan / deg : 0.0);
    f[14] = lg(maxlen) - lg(mean);
    f[15] = lg(span > 0 ? terms / span : 0.0);
}

// pbin of a row the network did not evaluate (no prediction could help it)
constexpr unsigned char PBIN_NONE = 255;

// Relative cost (in word operations) of counting a row of t terms with the
// accumulator a size bound of `cap` gets: a hash table of S slots is cleared
// and scanned (2S) and takes t probing inserts (2t); an LDS column bitmap of
// nwords words costs 2 nwords + t; a bitmap in global memory the same through
// L2 atomics, eight times slower.
__device__ __forceinline__ double count_cost(int64_t cap, int64_t t, int nwords, int lds_words,
                                             int hash_max, int dens)
{
    const int c = classify(cap, nwords, lds_words, hash_max, dens);
    if (c == 0) return 0.0;
    // a fixed part per row: the workgroup's start and the row's setup
    if (c < CLS_BITMAP) return 512.0 + 2.0 * (double) (64 << (c - 1)) + 2.0 * (double) t;
    if (c == CLS_BITMAP) return 512.0 + 2.0 * nwords + (double) t;
    return 8.0 * (512.0 + 2.0 * nwords + (double) t);
}

// Margins of the predicted sizing: a row is a candidate (features and a
// network evaluation) when a certain prediction could cut its count cost to
// PRED_CAND of the cost from its term bound, and is lowered when the expected
// cost of the lowered count plus a possible recount is below PRED_LOWER of it.
// Held-out graph products (R-MAT) showed the network over-confident far from
// its training families -- most lowered rows overflowed -- so a lowering must
// promise at least a halving, and rows promising less are not evaluated.
constexpr double PRED_CAND = 0.25;
constexpr double PRED_LOWER = 0.5;

// Rows a prediction could help: those where even a certain prediction of the
// smallest bin below the term bound would cost under PRED_CAND of the count
// from the term bound.  Only they get features and a network evaluation; one
// atomic per wavefront appends them to list.
__global__ void __launch_bounds__(256)
    k_pred_candidates(int64_t m, const int64_t *__restrict__ terms,
                      const int64_t *__restrict__ bound, int nwords, int lds_words,
                      int hash_max, int dens, int *__restrict__ list, int *__restrict__ count)
{
    const int64_t row = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    bool cand = false;
    if (row < m)
    {
        const int64_t t = terms[row], u = bound[row];
        if (u > mlp_bin_max(0))
            cand = count_cost(mlp_bin_max(0), t, nwords, lds_words, hash_max, dens) <
                   PRED_CAND * count_cost(u, t, nwords, lds_words, hash_max, dens);
    }
    const unsigned long long v = __ballot(cand);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0 && v) base = atomicAdd(count, (int) __popcll(v));
    base = __shfl(base, 0, 64);
    if (cand && list) list[base + __popcll(v & ((1ull << lane) - 1ull))] = (int) row;
}

// Count-pass size bound of every row from the network's bin probabilities
// (logits Z[row][6]): lowering the row's bound to the largest nnz of bin b
// costs the count with that bound's accumulator, plus, with probability
// P(nnz > bound_b) = 1 - (p_0 + ... + p_b), a second count from the term
// bound.  The bound is lowered to the bin of least expected cost when that
// cost is below PRED_LOWER of the count from the term bound; otherwise the term
// bound stays.  Uncertain rows (spread probabilities) and rows whose inserts
// dominate (hub rows) keep their term bound.  pbin gets the tau-rule bin (the
// smallest bin whose cumulative probability reaches tau) for the statistics.
__global__ void k_pred_cap(int64_t n, const int *__restrict__ list,
                           const int64_t *__restrict__ terms, const int64_t *__restrict__ ubound,
                           const float *__restrict__ F,
                           const float *__restrict__ Z, float tau, int nwords, int lds_words,
                           int hash_max, int dens, int64_t *__restrict__ cap,
                           unsigned char *__restrict__ pbin)
{
    // F[i] / Z[i] hold the features / logits of output row list[i]
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t row = list[i];
    const int64_t t = terms[row], u = ubound[row];
    // the longest A row among the row's terms puts that many distinct columns
    // in it (feature 2 is log2(1 + longest)): bins below it are impossible
    const double longest = rint(exp2((double) F[i * MLP_IN + 2]) - 1.0);
    const float *z = Z + i * MLP_OUT;
    float p[MLP_OUT];
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
    int tb = MLP_OUT - 1;
    float cum = 0.0f;
    for (int c = 0; c < MLP_OUT; c++)
    {
        p[c] = p[c] / s;
        cum = cum + p[c];
        if (cum >= tau && tb == MLP_OUT - 1 && c < MLP_OUT - 1) tb = c;
    }
    pbin[row] = (unsigned char) tb;
    int64_t c_out = u;
    if (u > 0)
    {
        const double e0 = count_cost(u, t, nwords, lds_words, hash_max, dens);
        double best = PRED_LOWER * e0;
        double below = 0.0;  // P(nnz <= bound_b)
        for (int b = 0; b < MLP_NBIN_BOUNDED; b++)
        {
            below += p[b];
            const int64_t bound = mlp_bin_max(b);
            if (bound >= u) break;  // the exact bound is already at most this bin's
            if ((double) bound < longest) continue;
            const double eb = count_cost(bound, t, nwords, lds_words, hash_max, dens) +
                              fmax(0.0, 1.0 - below) * e0;
            if (eb < best)
            {
                best = eb;
                c_out = bound;
            }
        }
    }
    cap[row] = c_out;
}

// Rows to count again from the exact bound (min(terms, span)): those whose
// count exceeded their predicted bound.  A hash table has at least twice its row's bound in slots
// and a failed insert leaves it full, so a count within the bound is exact.
__global__ void k_pred_recount(int64_t m, const int *__restrict__ rowcnt,
                               const int64_t *__restrict__ cap,
                               const int64_t *__restrict__ bound, int64_t *__restrict__ recount,
                               unsigned long long *__restrict__ nrecount)
{
    int64_t row = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    bool again = false;
    if (row < m)
    {
        const int64_t u = bound[row];
        again = cap[row] < u && (int64_t) rowcnt[row] > cap[row];
        recount[row] = again ? u : 0;
    }
    const unsigned long long v = __ballot(again);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(nrecount, (unsigned long long) __popcll(v));
}

// Over non-empty rows: [0] counted within the predicted bound, [1] counted
// again (the predicted bound was exceeded), [2] predicted bin differs from
// the bin of the exact nnz, [3] predicted bound not below the exact bound.
__global__ void __launch_bounds__(256)
    k_pred_stats(int64_t m, const int64_t *__restrict__ cap,
                 const int64_t *__restrict__ recount, const unsigned char *__restrict__ pbin,
                 const int64_t *__restrict__ bound, const int64_t *__restrict__ Cp,
                 unsigned long long *__restrict__ out)
{
    int64_t row = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    bool f[4] = {false, false, false, false};
    if (row < m && bound[row] > 0)
    {
        const bool lowered = cap[row] < bound[row];
        f[0] = lowered && recount[row] == 0;
        f[1] = lowered && recount[row] != 0;
        f[2] = pbin[row] != PBIN_NONE && pbin[row] != mlp_bin_of(Cp[row + 1] - Cp[row]);
        f[3] = !lowered;
    }
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        const unsigned long long v = __ballot(f[i]);
        if (lane == 0 && v) atomicAdd(&out[i], (unsigned long long) __popcll(v));
    }
}

// Term lists of one output row: entries' term counts, offsets, scatter, then
// each entry's list sorted by k (k is unique within an entry).  cols/off live
// in LDS (CAP > 0) or in global memory (CAP == 0: cols = Ci, off = gscratch).
// In global memory the counters are updated by atomics, which are performed
// in L2, so every other access to them is an agent-scope atomic load / store
// too and the phases are separated by device-scope fences: a plain store
// would otherwise only be visible to this CU, not to the atomics.
template <int CAP> __device__ __forceinline__ int off_ld(const int *off, int e)
{
    if (CAP > 0) return off[e];
    return __hip_atomic_load(off + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int CAP> __device__ __forceinline__ void off_st(int *off, int e, int v)
{
    if (CAP > 0) off[e] = v;
    else __hip_atomic_store(off + e, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int CAP> __device__ __forceinline__ void off_fence()
{
    if (CAP == 0) __threadfence();
}

template <int G, int CAP>
__global__ void __launch_bounds__(G)
    k_terms(const int *__restrict__ rows, int Lm, const int *__restrict__ Lp,
            const int *__restrict__ Li, const int *__restrict__ Lperm, int nL,
            const int *__restrict__ Ap, const int *__restrict__ Ai,
            const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
            const int64_t *__restrict__ toff, int64_t *__restrict__ ep, int *__restrict__ tA,
            int *__restrict__ tB, int *__restrict__ tK, int *__restrict__ tq,
            int *__restrict__ gscratch, int *__restrict__ err)
{
    // In global memory the row's columns are searched through a sample of
    // every ss-th column held in LDS: the sample's binary search picks a
    // segment of at most ss columns, searched in global memory (a few
    // accesses within one or two cache lines instead of ~log2(cnt) across
    // the row).
    constexpr int NSAMP = CAP > 0 ? 1 : 2048;
    __shared__ int s_cols[CAP > 0 ? CAP : 1];
    __shared__ int s_off[CAP > 0 ? CAP : 1];
    __shared__ int samp[NSAMP];
    __shared__ Expand<G> ex;
    __shared__ int wsum[G / 64];
    const int row = rows[blockIdx.x];
    const int64_t cbase = Cp[row];
    const int cnt = (int) (Cp[row + 1] - cbase);
    const int *cols;
    int *off;
    int ss = 1, ns = 0;
    if (CAP > 0)
    {
        for (int i = threadIdx.x; i < cnt; i += G) s_cols[i] = Ci[cbase + i];
        cols = s_cols;
        off = s_off;
    }
    else
    {
        cols = Ci + cbase;
        off = gscratch + cbase;
        while ((int64_t) ss * NSAMP < cnt) ss <<= 1;
        ns = (cnt + ss - 1) / ss;
        for (int i = threadIdx.x; i < ns; i += G) samp[i] = cols[i * ss];
    }
    // lower bound of col among the row's columns
    auto find = [&](int col) -> int {
        if (CAP > 0) return lower_bound_int(cols, cnt, col);
        // last sample <= col (the segment holding col, if any)
        const int i = max(0, lower_bound_int(samp, ns, col + 1) - 1);
        const int b = i * ss, e = min(cnt, b + ss);
        return b + lower_bound_int(cols + b, e - b, col);
    };
    for (int i = threadIdx.x; i < cnt; i += G) off_st<CAP>(off, i, 0);
    off_fence<CAP>();
    __syncthreads();
    RowCtx c = row_ctx(row, Lm, Lp, nL);
    for_each_term<G>(ex, c.q0, c.q1, c.koff, Li, Ap, [&](int t, int, int) {
        const int pos = find(Ai[t]);
        if (pos < cnt && cols[pos] == Ai[t]) atomicAdd(&off[pos], 1);
        else atomicOr(err, 1);
    });
    off_fence<CAP>();
    __syncthreads();
    // blocked exclusive scan of the per-entry counts
    const int64_t rbase = toff[row];
    {
        int per = (cnt + G - 1) / G;
        int e0 = min(cnt, (int) threadIdx.x * per), e1 = min(cnt, e0 + per);
        int s = 0;
        for (int e = e0; e < e1; e++) s += off_ld<CAP>(off, e);
        int tot;
        int pre = block_excl_scan_int<G>(s, &tot, wsum);
        for (int e = e0; e < e1; e++)
        {
            int v = off_ld<CAP>(off, e);
            off_st<CAP>(off, e, pre);
            ep[cbase + e] = rbase + pre;
            pre += v;
        }
        off_fence<CAP>();
        __syncthreads();
    }
    const int64_t rend = toff[row + 1];
    for_each_term<G>(ex, c.q0, c.q1, c.koff, Li, Ap, [&](int t, int q, int k) {
        const int col = Ai[t];
        int pos = find(col);
        if (pos >= cnt || cols[pos] != col)  // the row's columns miss a term: report, never write
        {
            atomicOr(err, 1);
            return;
        }
        int64_t slot = rbase + atomicAdd(&off[pos], 1);
        if (slot >= rend)
        {
            atomicOr(err, 1);
            return;
        }
        // tK holds the L index q until the lists are sorted (below)
        tA[slot] = t;
        tB[slot] = Lperm ? Lperm[q] : q;
        tK[slot] = q;
    });
    off_fence<CAP>();
    __threadfence_block();
    __syncthreads();
    // off[e] is now the end of entry e's list.  Sort each list by q, L's
    // storage order, which is the order the reference folds an entry's terms
    // in (gustavson.c:127-135 walks L's row ascending); q is unique within an
    // entry (A rows have distinct columns).  q order is ascending k whenever
    // L's rows are sorted; it also fixes the order for unsorted or repeated
    // row indices.  Then replace q by k.
    for (int e = threadIdx.x; e < cnt; e += G)
    {
        int64_t s1 = rbase + off_ld<CAP>(off, e);
        int64_t s0 = e == 0 ? rbase : rbase + off_ld<CAP>(off, e - 1);
        if (s0 < rbase || s1 > rend || s0 > s1)  // inconsistent counters: report, touch nothing
        {
            atomicOr(err, 1);
            continue;
        }
        for (int64_t i = s0 + 1; i < s1; i++)
        {
            int qq = tK[i], aa = tA[i], bb = tB[i];
            int64_t j = i - 1;
            while (j >= s0 && tK[j] > qq)
            {
                tK[j + 1] = tK[j];
                tA[j + 1] = tA[j];
                tB[j + 1] = tB[j];
                j--;
            }
            tK[j + 1] = qq;
            tA[j + 1] = aa;
            tB[j + 1] = bb;
        }
        if (tq)
            for (int64_t i = s0; i < s1; i++) tq[i] = tK[i];
        for (int64_t i = s0; i < s1; i++) tK[i] = Li[tK[i]] + c.koff;
    }
}

__global__ void k_set_last(int64_t *ep, int64_t n, int64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) ep[n] = v;
}

// ----------------------------------------------------------- numeric ----
__global__ void __launch_bounds__(256)
    k_num_pull(int64_t nnz, const int64_t *__restrict__ ep, const int *__restrict__ tA,
               const int *__restrict__ tB, const int *__restrict__ tK,
               const double *__restrict__ Lval, const double *__restrict__ d,
               const double *__restrict__ Ax, double *__restrict__ Cx)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    int64_t s0 = ep[e], s1 = ep[e + 1];
    double acc = 0.0;
    for (int64_t s = s0; s < s1; s++)
    {
        double v = Lval[tB[s]];
        if (d) v = v * d[tK[s]];
        acc = acc + v * Ax[tA[s]];
    }
    Cx[e] = acc;
}

// L entries [qr.x, qr.y) of the output rows holding the entries of 256-entry
// tile t (single-block L: consecutive rows own consecutive L entries).
__global__ void k_tile_qrange(int ntiles, int64_t nnz, int64_t m_out,
                              const int64_t *__restrict__ Cp, const int *__restrict__ Lp,
                              int2 *__restrict__ qr)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int64_t e0 = (int64_t) t * 256, e1 = min(nnz, e0 + 256) - 1;
    // row of entry e: the last r with Cp[r] <= e
    auto row_of = [&](int64_t e) {
        int64_t lo = 0, hi = m_out - 1;
        while (lo < hi)
        {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (Cp[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    const int64_t r0 = row_of(e0), r1 = row_of(e1);
    qr[t] = make_int2(Lp[r0], Lp[r1 + 1]);
}

// Tile refill of a single-block pull plan: one workgroup per 256 entries.
// The scale factor of every L entry q of the tile's rows, Lval[Lperm[q]] *
// d[Li[q]], is formed once into LDS (when there are at most TS_CAP of them;
// else per term); each entry then folds its terms in order, scale(tq) *
// Ax[tA] from +0.0 -- k_num_pull's products and fold, with one index and one
// A value read per term instead of three indices and three gathers.
// BV (bound values, the plan's tile layout): Lval is the bound L-order value
// array lv (lv[q] = Lval[Lperm[q]]) and Ax the bound term-order A values ta
// (ta[s] = Ax[tA[s]]), both read without an index.
constexpr int TS_CAP = 1024;
template <bool BV>
__global__ void __launch_bounds__(256)
    k_num_pull_tile(int64_t nnz, const int2 *__restrict__ tqr, const int64_t *__restrict__ ep,
                    const int *__restrict__ tq, const int *__restrict__ tA,
                    const int *__restrict__ Li, const int *__restrict__ Lperm,
                    const double *__restrict__ Lval, const double *__restrict__ d,
                    const double *__restrict__ Ax, double *__restrict__ Cx)
{
    __shared__ double sc[TS_CAP];
    const int2 qr = tqr[blockIdx.x];
    const int nq = qr.y - qr.x;
    const bool table = nq <= TS_CAP;  // the same for the whole workgroup
    const int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x;
    const bool live = e < nnz;
    // this entry's term range, loaded before the table is built
    const int64_t s0 = live ? ep[e] : 0, s1 = live ? ep[e + 1] : 0;
    if (table)
    {
        for (int j = threadIdx.x; j < nq; j += 256)
        {
            const int q = qr.x + j;
            double v = Lval[(!BV && Lperm) ? Lperm[q] : q];
            if (d) v = v * d[Li[q]];
            sc[j] = v;
        }
    }
    __syncthreads();
    if (!live) return;
    double acc = 0.0;
    for (int64_t s = s0; s < s1; s++)
    {
        const int q = tq[s];
        double v;
        if (table) v = sc[q - qr.x];
        else
        {
            v = Lval[(!BV && Lperm) ? Lperm[q] : q];
            if (d) v = v * d[Li[q]];
        }
        acc = acc + v * Ax[BV ? s : tA[s]];
    }
    Cx[e] = acc;
}

// Several single refills (products of single-block pull plans) in one launch:
// the tiles of node 0, then node 1, ...; each workgroup finds its node from the
// tile offsets and runs k_num_pull_tile's body on it.  The descriptors travel
// by value in the kernel arguments.
struct TileNode
{
    const int2 *tqr;
    const int64_t *ep;
    const int *tq, *tA, *Li, *Lperm;
    const double *Lval, *d, *Ax;
    double *Cx;
    int64_t nnz;
    int tile0;
};
constexpr int TN_MAX = 16;
struct TileNodes
{
    int n;
    TileNode nd[TN_MAX];
};

__global__ void __launch_bounds__(256) k_num_pull_tile_multi(const TileNodes N)
{
    __shared__ double sc[TS_CAP];
    int j = 0;
    for (int t = 1; t < N.n; t++)
        if (N.nd[t].tile0 <= (int) blockIdx.x) j = t;
    const TileNode &D = N.nd[j];
    const int tile = (int) blockIdx.x - D.tile0;
    const int2 qr = D.tqr[tile];
    const int nq = qr.y - qr.x;
    const bool table = nq <= TS_CAP;  // the same for the whole workgroup
    const int64_t e = (int64_t) tile * 256 + threadIdx.x;
    const bool live = e < D.nnz;
    const int64_t s0 = live ? D.ep[e] : 0, s1 = live ? D.ep[e + 1] : 0;
    if (table)
    {
        for (int jj = threadIdx.x; jj < nq; jj += 256)
        {
            const int q = qr.x + jj;
            double v = D.Lval[D.Lperm ? D.Lperm[q] : q];
            if (D.d) v = v * D.d[D.Li[q]];
            sc[jj] = v;
        }
    }
    __syncthreads();
    if (!live) return;
    double acc = 0.0;
    for (int64_t s = s0; s < s1; s++)
    {
        const int q = D.tq[s];
        double v;
        if (table) v = sc[q - qr.x];
        else
        {
            v = D.Lval[D.Lperm ? D.Lperm[q] : q];
            if (D.d) v = v * D.d[D.Li[q]];
        }
        acc = acc + v * D.Ax[D.tA[s]];
    }
    D.Cx[e] = acc;
}

// Bound tile layout: ta[s] = Ax[tA[s]] (term order), lv[q] = Lval[Lperm[q]]
// (L order).
__global__ void __launch_bounds__(256)
    k_bind_tile(int64_t terms, int64_t nl, const int *__restrict__ tA,
                const int *__restrict__ Lperm, const double *__restrict__ Lval,
                const double *__restrict__ Ax, double *__restrict__ ta, double *__restrict__ lv)
{
    const int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
    if (i < terms) ta[i] = Ax[tA[i]];
    if (i < nl) lv[i] = Lval[Lperm ? Lperm[i] : i];
}

constexpr int NT_REG = 4;
typedef double dvec2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256)
    k_num_pull_batched(int64_t nnz, const int64_t *__restrict__ ep, const int *__restrict__ tA,
                       const int *__restrict__ tB, const int *__restrict__ tK,
                       const double *__restrict__ Lval, const double *__restrict__ D,
                       int64_t ldd, int K, const double *__restrict__ Ax,
                       double *__restrict__ C, int64_t ldc)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const int64_t s0 = ep[e];
    const int nt = (int) (ep[e + 1] - s0);
    double bv[NT_REG], av[NT_REG];
    int kk[NT_REG];
#pragma unroll
    for (int j = 0; j < NT_REG; j++)
    {
        if (j < nt)
        {
            bv[j] = Lval[tB[s0 + j]];
            av[j] = Ax[tA[s0 + j]];
            kk[j] = tK[s0 + j];
        }
        else
        {
            bv[j] = 0.0;
            av[j] = 0.0;
            kk[j] = 0;
        }
    }
    if (nt <= NT_REG)
    {
        for (int r = 0; r < K; r++)
        {
            const double *dr = D + (int64_t) r * ldd;
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < NT_REG; j++)
                if (j < nt) acc = acc + (bv[j] * dr[kk[j]]) * av[j];
            __builtin_nontemporal_store(acc, C + (int64_t) r * ldc + e);
        }
    }
    else
    {
        for (int r = 0; r < K; r++)
        {
            const double *dr = D + (int64_t) r * ldd;
            double acc = 0.0;
            for (int j = 0; j < nt; j++)
            {
                int64_t s = s0 + j;
                acc = acc + (Lval[tB[s]] * dr[tK[s]]) * Ax[tA[s]];
            }
            __builtin_nontemporal_store(acc, C + (int64_t) r * ldc + e);
        }
    }
}

// Batched pull refill, r-chunked: grid covers (entry tile, chunk of RC
// refills).  Tiles map XCD-aware: the chunks of one tile run back to back on
// one XCD, so its term data is read from HBM once and then hit in that XCD's
// L2.  EPT consecutive entries per thread (EPT = 2 stores 16 bytes per lane).
template <int EPT, int RC>
__global__ void __launch_bounds__(256)
    k_num_pull_rchunk(int64_t nnz, int ntiles, int nchunks, const int64_t *__restrict__ ep,
                      const int *__restrict__ tA, const int *__restrict__ tB,
                      const int *__restrict__ tK, const double *__restrict__ Lval,
                      const double *__restrict__ D, int64_t ldd, int K,
                      const double *__restrict__ Ax, double *__restrict__ C, int64_t ldc)
{
    const int b = blockIdx.x;
    const int xcd = b & 7, s = b >> 3;
    const int tile = (s / nchunks) * 8 + xcd;
    const int chunk = s % nchunks;
    if (tile >= ntiles) return;
    const int64_t e0 = ((int64_t) tile * blockDim.x + threadIdx.x) * EPT;
    if (e0 >= nnz) return;
    const int r0 = chunk * RC;
    const int rn = min(RC, K - r0);
    double bv[EPT][2], av[EPT][2];
    int kk[EPT][2], nt[EPT];
    bool simple = true;
#pragma unroll
    for (int u = 0; u < EPT; u++)
    {
        int64_t e = e0 + u;
        nt[u] = 0;
        if (e < nnz)
        {
            int64_t sa = ep[e];
            nt[u] = (int) (ep[e + 1] - sa);
            if (nt[u] > 2) simple = false;
#pragma unroll
            for (int j = 0; j < 2; j++)
            {
                if (j < nt[u])
                {
                    bv[u][j] = Lval[tB[sa + j]];
                    av[u][j] = Ax[tA[sa + j]];
                    kk[u][j] = tK[sa + j];
                }
                else
                {
                    bv[u][j] = 0.0;
                    av[u][j] = 0.0;
                    kk[u][j] = 0;
                }
            }
        }
    }
    const bool full = (e0 + EPT <= nnz);
    if (simple && rn == RC)
    {
        double dv[RC][EPT][2];
#pragma unroll
        for (int r = 0; r < RC; r++)
        {
            const double *dr = D + (int64_t) (r0 + r) * ldd;
#pragma unroll
            for (int u = 0; u < EPT; u++)
#pragma unroll
                for (int j = 0; j < 2; j++) dv[r][u][j] = (j < nt[u]) ? dr[kk[u][j]] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < RC; r++)
        {
            double out[EPT];
#pragma unroll
            for (int u = 0; u < EPT; u++)
            {
                double acc = 0.0;
                if (nt[u] > 0) acc = acc + (bv[u][0] * dv[r][u][0]) * av[u][0];
                if (nt[u] > 1) acc = acc + (bv[u][1] * dv[r][u][1]) * av[u][1];
                out[u] = acc;
            }
            double *cr = C + (int64_t) (r0 + r) * ldc + e0;
            if (EPT == 2 && full && ((ldc & 1) == 0))
            {
                dvec2 v2;
                v2.x = out[0];
                v2.y = out[EPT - 1];
                __builtin_nontemporal_store(v2, (dvec2 *) cr);
            }
            else
            {
#pragma unroll
                for (int u = 0; u < EPT; u++)
                    if (e0 + u < nnz) __builtin_nontemporal_store(out[u], cr + u);
            }
        }
        return;
    }
    // general lists / partial chunk
    for (int r = r0; r < r0 + rn; r++)
    {
        const double *dr = D + (int64_t) r * ldd;
#pragma unroll
        for (int u = 0; u < EPT; u++)
        {
            int64_t e = e0 + u;
            if (e >= nnz) break;
            int64_t sa = ep[e];
            double acc = 0.0;
            for (int j = 0; j < nt[u]; j++)
            {
                int64_t q = sa + j;
                acc = acc + (Lval[tB[q]] * dr[tK[q]]) * Ax[tA[q]];
            }
            __builtin_nontemporal_store(acc, C + (int64_t) r * ldc + e);
        }
    }
}

// Batch preparation for the interleaved refill.
// k_gather_terms: each term's operand pair tv[s] = (L value, A value) in term
// order (depends only on the operand values).
// d transposition: D ([K][ldd], refill-major) -> DT ([m][ldt], k-major: the K
// values of one k contiguous), through a 16 x 64 LDS tile (k_prep_batch).
constexpr int TR_K = 64;   // k per transpose tile
constexpr int TR_R = 16;   // refills per transpose tile

__global__ void __launch_bounds__(256)
    k_gather_terms(int64_t terms, const int *__restrict__ tA, const int *__restrict__ tB,
                   const double *__restrict__ Lval, const double *__restrict__ Ax,
                   double *__restrict__ tv)
{
    int64_t s = (int64_t) blockIdx.x * 256 + threadIdx.x;
    if (s >= terms) return;
    dvec2 v;
    v.x = Lval[tB[s]];
    v.y = Ax[tA[s]];
    *(dvec2 *) (tv + 2 * s) = v;
}

// Batch preparation in one launch: workgroups [0, ngather) gather operand
// pairs (as k_gather_terms), the rest transpose d tiles; the two are
// independent, so they run side by side.
__global__ void __launch_bounds__(256)
    k_prep_batch(int ngather, int64_t terms, const int *__restrict__ tA,
                 const int *__restrict__ tB, const double *__restrict__ Lval,
                 const double *__restrict__ Ax, double *__restrict__ tv, int m, int K,
                 const double *__restrict__ D, int64_t ldd, double *__restrict__ DT, int ldt)
{
    __shared__ double tile[TR_R][TR_K + 1];
    if ((int) blockIdx.x < ngather)
    {
        const int64_t s = (int64_t) blockIdx.x * 256 + threadIdx.x;
        if (s >= terms) return;
        dvec2 v;
        v.x = Lval[tB[s]];
        v.y = Ax[tA[s]];
        *(dvec2 *) (tv + 2 * s) = v;
        return;
    }
    const int bt = (int) blockIdx.x - ngather;
    const int nkt = (m + TR_K - 1) / TR_K;
    const int kt = bt % nkt, rt = bt / nkt;
    const int k0 = kt * TR_K, r0 = rt * TR_R;
    for (int i = threadIdx.x; i < TR_R * TR_K; i += 256)
    {
        int rr = i / TR_K, kk = i % TR_K;
        int r = r0 + rr, k = k0 + kk;
        tile[rr][kk] = (r < K && k < m) ? D[(int64_t) r * ldd + k] : 0.0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TR_R * TR_K; i += 256)
    {
        int kk = i / TR_R, rr = i % TR_R;
        int r = r0 + rr, k = k0 + kk;
        if (k < m && r < ldt) DT[(int64_t) k * ldt + r] = tile[rr][kk];
    }
}

// Interleaved batched refill: grid over (entry tile, chunk of RC refills),
// XCD-aware tile mapping; each term reads its RC values of d as one
// contiguous, aligned run of DT.
template <int RC>
__global__ void __launch_bounds__(256)
    k_num_pull_interleaved(int64_t nnz, int ntiles, int nchunks, const int64_t *__restrict__ ep,
                           const int *__restrict__ tK, const double *__restrict__ tv,
                           const double *__restrict__ DT, int ldt, int K,
                           double *__restrict__ C, int64_t ldc)
{
    const int b = blockIdx.x;
    const int xcd = b & 7, sl = b >> 3;
    const int tile = (sl / nchunks) * 8 + xcd;
    const int chunk = sl % nchunks;
    if (tile >= ntiles) return;
    const int64_t e = (int64_t) tile * 256 + threadIdx.x;
    if (e >= nnz) return;
    const int r0 = chunk * RC;
    const int64_t s0 = ep[e], s1 = ep[e + 1];
    double out[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) out[r] = 0.0;
    for (int64_t s = s0; s < s1; s++)
    {
        const int k = tK[s];
        const dvec2 ba = *(const dvec2 *) (tv + 2 * s);
        const dvec2 *dk = (const dvec2 *) (DT + (int64_t) k * ldt + r0);
        dvec2 dv[RC / 2];
#pragma unroll
        for (int h = 0; h < RC / 2; h++) dv[h] = dk[h];
#pragma unroll
        for (int h = 0; h < RC / 2; h++)
        {
            out[2 * h] = out[2 * h] + (ba.x * dv[h].x) * ba.y;
            out[2 * h + 1] = out[2 * h + 1] + (ba.x * dv[h].y) * ba.y;
        }
    }
    const int rn = min(RC, K - r0);
    double *cp = C + (int64_t) r0 * ldc + e;
    if (rn == RC)
    {
#pragma unroll
        for (int r = 0; r < RC; r++) __builtin_nontemporal_store(out[r], cp + (int64_t) r * ldc);
    }
    else
    {
        for (int r = 0; r < rn; r++) __builtin_nontemporal_store(out[r], cp + (int64_t) r * ldc);
    }
}

// The interleaved refill reading each term's operands through tA / tB (the
// value arrays stay in L2) instead of a pre-gathered pair array: no gather
// pass per batch, 8 instead of 16 bytes of term data per term.
template <int RC>
__global__ void __launch_bounds__(256)
    k_num_pull_interleaved_g(int64_t nnz, int ntiles, int nchunks, const int64_t *__restrict__ ep,
                             const int *__restrict__ tA, const int *__restrict__ tB,
                             const int *__restrict__ tK, const double *__restrict__ Lval,
                             const double *__restrict__ Ax, const double *__restrict__ DT,
                             int ldt, int K, double *__restrict__ C, int64_t ldc)
{
    const int b = blockIdx.x;
    const int xcd = b & 7, sl = b >> 3;
    const int tile = (sl / nchunks) * 8 + xcd;
    const int chunk = sl % nchunks;
    if (tile >= ntiles) return;
    const int64_t e = (int64_t) tile * 256 + threadIdx.x;
    if (e >= nnz) return;
    const int r0 = chunk * RC;
    const int64_t s0 = ep[e], s1 = ep[e + 1];
    double out[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) out[r] = 0.0;
    for (int64_t s = s0; s < s1; s++)
    {
        const int k = tK[s];
        const double bx = Lval[tB[s]], ay = Ax[tA[s]];
        const dvec2 *dk = (const dvec2 *) (DT + (int64_t) k * ldt + r0);
        dvec2 dv[RC / 2];
#pragma unroll
        for (int h = 0; h < RC / 2; h++) dv[h] = dk[h];
#pragma unroll
        for (int h = 0; h < RC / 2; h++)
        {
            out[2 * h] = out[2 * h] + (bx * dv[h].x) * ay;
            out[2 * h + 1] = out[2 * h + 1] + (bx * dv[h].y) * ay;
        }
    }
    const int rn = min(RC, K - r0);
    double *cp = C + (int64_t) r0 * ldc + e;
    if (rn == RC)
    {
#pragma unroll
        for (int r = 0; r < RC; r++) __builtin_nontemporal_store(out[r], cp + (int64_t) r * ldc);
    }
    else
    {
        for (int r = 0; r < rn; r++) __builtin_nontemporal_store(out[r], cp + (int64_t) r * ldc);
    }
}

// One refill over the pre-gathered operand pairs: C[e] = sum over the entry's
// terms (ascending k) of (L value * d[k]) * A value -- the fold of every
// other refill kernel.  One 256-entry tile per workgroup (with a single
// refill no tile's data is reused by another workgroup).
__global__ void __launch_bounds__(256)
    k_num_pull_tv1(int64_t nnz, const int64_t *__restrict__ ep, const int *__restrict__ tK,
                   const double *__restrict__ tv, const double *__restrict__ d,
                   double *__restrict__ C)
{
    const int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x;
    if (e >= nnz) return;
    const int64_t s0 = ep[e], s1 = ep[e + 1];
    double out = 0.0;
    for (int64_t s = s0; s < s1; s++)
    {
        const dvec2 ba = *(const dvec2 *) (tv + 2 * s);
        out = out + (ba.x * d[tK[s]]) * ba.y;
    }
    __builtin_nontemporal_store(out, C + e);
}

// Tile form of the interleaved batched refill (variant 7): grid over (entry
// tile, chunk of RC refills) as k_num_pull_interleaved.  The workgroup first
// forms, for every L entry q of its tile's rows, the RC scale factors
// Lval[Lperm[q]] * d_r[Li[q]] (one aligned run of DT per L entry) into LDS;
// each term then reads its L index and its A value and adds scale * A value
// for each of the RC refills -- the products and the fold of
// k_num_pull_interleaved, with d read per L entry rather than per term and no
// operand-pair array.  Tiles of more than TB_CAP L entries form each term's
// factors from DT directly.
constexpr int TB_CAP = 256;
template <int RC, bool BV>
__global__ void __launch_bounds__(256)
    k_num_pull_tile_batched(int64_t nnz, int ntiles, int nchunks, const int2 *__restrict__ tqr,
                            const int64_t *__restrict__ ep, const int *__restrict__ tq,
                            const int *__restrict__ tA, const int *__restrict__ Li,
                            const int *__restrict__ Lperm, const double *__restrict__ Lval,
                            const double *__restrict__ Ax, const double *__restrict__ DT, int ldt,
                            int K, double *__restrict__ C, int64_t ldc)
{
    __shared__ dvec2 tab[TB_CAP * (RC / 2)];  // tab[j * RC/2 + h]: refills r0+2h, r0+2h+1
    const int b = blockIdx.x;
    const int xcd = b & 7, sl = b >> 3;
    const int tile = (sl / nchunks) * 8 + xcd;
    const int chunk = sl % nchunks;
    if (tile >= ntiles) return;  // the whole workgroup
    const int r0 = chunk * RC;
    const int2 qr = tqr[tile];
    const int nq = qr.y - qr.x;
    const bool table = nq <= TB_CAP;  // the same for the whole workgroup
    const int64_t e = (int64_t) tile * 256 + threadIdx.x;
    const bool live = e < nnz;
    const int64_t s0 = live ? ep[e] : 0, s1 = live ? ep[e + 1] : 0;
    if (table)
    {
        for (int j = threadIdx.x; j < nq; j += 256)
        {
            const int q = qr.x + j;
            const double lv = Lval[(!BV && Lperm) ? Lperm[q] : q];
            const dvec2 *dk = (const dvec2 *) (DT + (int64_t) Li[q] * ldt + r0);
#pragma unroll
            for (int h = 0; h < RC / 2; h++)
            {
                const dvec2 dv = dk[h];
                dvec2 t;
                t.x = lv * dv.x;
                t.y = lv * dv.y;
                tab[j * (RC / 2) + h] = t;
            }
        }
    }
    __syncthreads();
    if (!live) return;
    double out[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) out[r] = 0.0;
    for (int64_t s = s0; s < s1; s++)
    {
        const int q = tq[s];
        const double a = Ax[BV ? s : tA[s]];
        if (table)
        {
            const dvec2 *t = tab + (q - qr.x) * (RC / 2);
#pragma unroll
            for (int h = 0; h < RC / 2; h++)
            {
                const dvec2 f = t[h];
                out[2 * h] = out[2 * h] + f.x * a;
                out[2 * h + 1] = out[2 * h + 1] + f.y * a;
            }
        }
        else
        {
            const double lv = Lval[(!BV && Lperm) ? Lperm[q] : q];
            const dvec2 *dk = (const dvec2 *) (DT + (int64_t) Li[q] * ldt + r0);
#pragma unroll
            for (int h = 0; h < RC / 2; h++)
            {
                const dvec2 dv = dk[h];
                out[2 * h] = out[2 * h] + (lv * dv.x) * a;
                out[2 * h + 1] = out[2 * h + 1] + (lv * dv.y) * a;
            }
        }
    }
    const int rn = min(RC, K - r0);
    double *cp = C + (int64_t) r0 * ldc + e;
    if (rn == RC)
    {
#pragma unroll
        for (int r = 0; r < RC; r++) __builtin_nontemporal_store(out[r], cp + (int64_t) r * ldc);
    }
    else
    {
        for (int r = 0; r < rn; r++) __builtin_nontemporal_store(out[r], cp + (int64_t) r * ldc);
    }
}

typedef double dvec4 __attribute__((ext_vector_type(4)));

// Interleaved batched refill with EPT adjacent entries per thread: one
// EPT*8-byte store per refill (VEC: C rows are EPT-aligned).
template <int RC, int EPT, bool VEC>
__global__ void __launch_bounds__(256)
    k_num_pull_interleaved_v(int64_t nnz, int ntiles, int nchunks,
                             const int64_t *__restrict__ ep, const int *__restrict__ tK,
                             const double *__restrict__ tv, const double *__restrict__ DT,
                             int ldt, int K, double *__restrict__ C, int64_t ldc)
{
    const int b = blockIdx.x;
    const int xcd = b & 7, sl = b >> 3;
    const int tile = (sl / nchunks) * 8 + xcd;
    const int chunk = sl % nchunks;
    if (tile >= ntiles) return;
    const int64_t e0 = ((int64_t) tile * 256 + threadIdx.x) * EPT;
    if (e0 >= nnz) return;
    const int r0 = chunk * RC;
    int64_t sb[EPT + 1];
#pragma unroll
    for (int u = 0; u <= EPT; u++) sb[u] = ep[min(e0 + u, nnz)];
    double out[EPT][RC];
#pragma unroll
    for (int u = 0; u < EPT; u++)
#pragma unroll
        for (int r = 0; r < RC; r++) out[u][r] = 0.0;
#pragma unroll
    for (int u = 0; u < EPT; u++)
    {
        for (int64_t s = sb[u]; s < sb[u + 1]; s++)
        {
            const int k = tK[s];
            const dvec2 ba = *(const dvec2 *) (tv + 2 * s);
            const dvec2 *dk = (const dvec2 *) (DT + (int64_t) k * ldt + r0);
            dvec2 dv[RC / 2];
#pragma unroll
            for (int h = 0; h < RC / 2; h++) dv[h] = dk[h];
#pragma unroll
            for (int h = 0; h < RC / 2; h++)
            {
                out[u][2 * h] = out[u][2 * h] + (ba.x * dv[h].x) * ba.y;
                out[u][2 * h + 1] = out[u][2 * h + 1] + (ba.x * dv[h].y) * ba.y;
            }
        }
    }
    const int rn = min(RC, K - r0);
    double *cp = C + (int64_t) r0 * ldc + e0;
    const bool full = e0 + EPT <= nnz;
    if (VEC && full && rn == RC)
    {
#pragma unroll
        for (int r = 0; r < RC; r++)
        {
            if (EPT == 4)
            {
                dvec4 v;
                v.x = out[0][r];
                v.y = out[1 % EPT][r];
                v.z = out[2 % EPT][r];
                v.w = out[3 % EPT][r];
                __builtin_nontemporal_store(v, (dvec4 *) (cp + (int64_t) r * ldc));
            }
            else if (EPT == 2)
            {
                dvec2 v;
                v.x = out[0][r];
                v.y = out[1 % EPT][r];
                __builtin_nontemporal_store(v, (dvec2 *) (cp + (int64_t) r * ldc));
            }
            else
                __builtin_nontemporal_store(out[0][r], cp + (int64_t) r * ldc);
        }
    }
    else
    {
        for (int r = 0; r < rn; r++)
#pragma unroll
            for (int u = 0; u < EPT; u++)
                if (e0 + u < nnz) __builtin_nontemporal_store(out[u][r], cp + (int64_t) r * ldc + u);
    }
}

// Ordered row accumulator: k walks L's row in storage order (one barrier per
// k); the columns of one A row are distinct, so its terms update distinct
// slots in parallel.  CAP == 0: values accumulate in Cx itself.
template <int G, int CAP>
__global__ void __launch_bounds__(G)
    k_num_rowacc(const int *__restrict__ rows, int Lm, const int *__restrict__ Lp,
                 const int *__restrict__ Li, const int *__restrict__ Lperm, int nL,
                 const int *__restrict__ Ap, const int *__restrict__ Ai,
                 const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
                 const double *__restrict__ Lval, const double *__restrict__ d,
                 const double *__restrict__ Ax, double *__restrict__ Cx)
{
    __shared__ int s_cols[CAP > 0 ? CAP : 1];
    __shared__ double s_val[CAP > 0 ? CAP : 1];
    const int row = rows[blockIdx.x];
    const int64_t cbase = Cp[row];
    const int cnt = (int) (Cp[row + 1] - cbase);
    const int *cols = CAP > 0 ? s_cols : Ci + cbase;
    double *val = CAP > 0 ? s_val : Cx + cbase;
    for (int i = threadIdx.x; i < cnt; i += G)
    {
        if (CAP > 0) s_cols[i] = Ci[cbase + i];
        val[i] = 0.0;
    }
    if (CAP == 0) __threadfence_block();
    __syncthreads();
    RowCtx c = row_ctx(row, Lm, Lp, nL);
    // L entries go G at a time: their A-row ranges and scale factors are
    // loaded together into LDS, and each A row's first G terms are loaded
    // before the previous row's updates (unconditionally: index 0 stands in
    // for a lane without a term), so the dependent loads of consecutive k
    // overlap the barrier between them.  Updates stay in ascending k.
    __shared__ int m_a[G], m_b[G];
    __shared__ double m_v[G];
    for (int qb = c.q0; qb < c.q1; qb += G)
    {
        const int nqc = min(G, c.q1 - qb);
        if ((int) threadIdx.x < nqc)
        {
            const int q = qb + (int) threadIdx.x;
            const int k = Li[q] + c.koff;
            double v = Lval[Lperm ? Lperm[q] : q];
            if (d) v = v * d[k];
            m_a[threadIdx.x] = Ap[k];
            m_b[threadIdx.x] = Ap[k + 1];
            m_v[threadIdx.x] = v;
        }
        __syncthreads();
        int tc = m_a[0] + (int) threadIdx.x;
        bool hc = tc < m_b[0];
        int colc = Ai[hc ? tc : 0];
        double xc = Ax[hc ? tc : 0];
        for (int j = 0; j < nqc; j++)
        {
            const int b = m_b[j];
            const double v = m_v[j];
            const bool more = j + 1 < nqc;
            const int tn = (more ? m_a[j + 1] : 0) + (int) threadIdx.x;
            const bool hn = more && tn < m_b[j + 1];
            const int coln = Ai[hn ? tn : 0];
            const double xn = Ax[hn ? tn : 0];
            if (hc)
            {
                const int pos = lower_bound_int(cols, cnt, colc);
                val[pos] = val[pos] + v * xc;
            }
            for (int t = tc + G; t < b; t += G)  // A rows longer than G
            {
                const int pos = lower_bound_int(cols, cnt, Ai[t]);
                val[pos] = val[pos] + v * Ax[t];
            }
            if (CAP == 0) __threadfence_block();
            __syncthreads();
            tc = tn;
            hc = hn;
            colc = coln;
            xc = xn;
        }
    }
    if (CAP > 0)
        for (int i = threadIdx.x; i < cnt; i += G) Cx[cbase + i] = s_val[i];
}

// Wavefront class of a row-accumulator row (ids[i]) of at most 1024 entries:
// 2 / 3 / 4 for a column span of at most 256 / 512 / 1024 (the window
// accumulator), else 5 / 6 / 7 for at most 256 / 512 / 1024 entries (the
// accumulator over the row's columns with a search); 1 for a longer row (the
// workgroup accumulator).  Rows hold at least one entry, in ascending columns.
constexpr int RW_CLASSES = 8;
__global__ void k_rowacc_span_class(int n, const int *__restrict__ ids,
                                    const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
                                    unsigned char *__restrict__ cls, int *__restrict__ counts)
{
    __shared__ int hist[RW_CLASSES];
    if (threadIdx.x < RW_CLASSES) hist[threadIdx.x] = 0;
    __syncthreads();
    const int i = (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
    {
        const int row = ids[i];
        const int64_t b = Cp[row], e = Cp[row + 1];
        const int span = Ci[e - 1] - Ci[b] + 1;
        const int64_t cnt = e - b;
        const int c = span <= 256 ? 2 : span <= 512 ? 3 : span <= 1024 ? 4
                    : cnt <= 256  ? 5 : cnt <= 512  ? 6 : cnt <= 1024 ? 7 : 1;
        cls[i] = (unsigned char) c;
        atomicAdd(&hist[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < RW_CLASSES && hist[threadIdx.x])
        atomicAdd(&counts[threadIdx.x], hist[threadIdx.x]);
}

// Window class of a rank-bin row (ids[i]): 2 / 3 / 4 for a column span of
// at most 2048 / 4096 / 8192, else 1.
__global__ void k_rowacc_wide_class(int n, const int *__restrict__ ids,
                                    const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
                                    unsigned char *__restrict__ cls, int *__restrict__ counts)
{
    __shared__ int hist[5];
    if (threadIdx.x < 5) hist[threadIdx.x] = 0;
    __syncthreads();
    const int i = (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
    {
        const int row = ids[i];
        const int64_t b = Cp[row], e = Cp[row + 1];
        const int span = Ci[e - 1] - Ci[b] + 1;
        const int c = span <= 2048 ? 2 : span <= 4096 ? 3 : span <= 8192 ? 4 : 1;
        cls[i] = (unsigned char) c;
        atomicAdd(&hist[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < 5 && hist[threadIdx.x]) atomicAdd(&counts[threadIdx.x], hist[threadIdx.x]);
}

// Ordered row accumulator, one wavefront per row, for rows whose columns span
// at most WS: a dense LDS accumulator over [lo, lo + WS) (lo = the row's first
// column) indexed by column - lo, so a term needs no column search.  k walks
// L's row in storage order; the columns of one A row are distinct, so one
// A row's terms update distinct slots together, and the wavefront's LDS
// accesses complete in issue order, so consecutive k update a slot in
// ascending k with no barrier.  The row is then gathered through its columns.
// Descriptors of 64 L entries at a time are loaded one per lane and read with
// readlane; the first U * 64 terms of each A row are loaded before the
// previous row's updates (index 0 stands in for a slot without a term:
// unconditional loads measured faster than masked ones unless whole chunks
// are empty, so U covers the mean A row).
//   SEARCH: rows of at most WS entries whose span exceeds the windows: the
//   accumulator holds the row's entries in order, the row's columns staged in
//   LDS, and a term's slot is found by binary search among them.
//   NWV: wavefronts (rows) per workgroup; the 2048- and 4096-column windows
//   of rank-bin rows run one per workgroup.
template <int WS, int U, bool SEARCH = false, int NWV = 4>
__global__ void __launch_bounds__(64 * NWV)
    k_num_rowacc_win(const int *__restrict__ rows, int nrows, int Lm, const int *__restrict__ Lp,
                     const int *__restrict__ Li, const int *__restrict__ Lperm, int nL,
                     const int *__restrict__ Ap, const int *__restrict__ Ai,
                     const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
                     const double *__restrict__ Lval, const double *__restrict__ d,
                     const double *__restrict__ Ax, double *__restrict__ Cx)
{
    // windows past 32 KB (8192 columns) live in dynamic LDS (64 KB, past
    // the static limit)
    constexpr bool DYN = (size_t) NWV * WS * sizeof(double) > 32768;
    __shared__ double acc[DYN ? 1 : NWV][DYN ? 1 : WS];
    extern __shared__ double dacc[];
    __shared__ int scol[SEARCH ? NWV : 1][SEARCH ? WS : 1];
    const int w = (int) threadIdx.x >> 6, lane = (int) threadIdx.x & 63;
    const int ri = (int) blockIdx.x * NWV + w;
    if (ri >= nrows) return;  // whole wavefronts; no workgroup barrier follows
    double *a = DYN ? dacc + (size_t) w * WS : acc[DYN ? 0 : w];
    const int row = rows[ri];
    const int64_t cb = Cp[row];
    const int cnt = (int) (Cp[row + 1] - cb);
    const int lo = Ci[cb];
    const int ncol = min(cnt, WS);
    int *sc = scol[SEARCH ? w : 0];
    if (SEARCH)
        for (int i = lane; i < ncol; i += 64) sc[i] = Ci[cb + i];
    // slot of a column: its offset in the window, or its position in the row
    auto slot = [&](int col) -> unsigned {
        if (!SEARCH) return (unsigned) (col - lo);
        const int p = lower_bound_int(sc, ncol, col);
        return p < ncol && sc[p] == col ? (unsigned) p : (unsigned) WS;
    };
    for (int i = lane; i < WS; i += 64) a[i] = 0.0;
    const RowCtx c = row_ctx(row, Lm, Lp, nL);
    for (int qb = c.q0; qb < c.q1; qb += 64)
    {
        const int nq = min(64, c.q1 - qb);
        int ma = 0, mb = 0;
        double mv = 0.0;
        if (lane < nq)
        {
            const int q = qb + lane;
            const int k = Li[q] + c.koff;
            double v = Lval[Lperm ? Lperm[q] : q];
            if (d) v = v * d[k];
            ma = Ap[k];
            mb = Ap[k + 1];
            mv = v;
        }
        const int mvlo = __double2loint(mv), mvhi = __double2hiint(mv);
        int ta = __builtin_amdgcn_readlane(ma, 0), tb = __builtin_amdgcn_readlane(mb, 0);
        int colc[U];
        double xc[U];
#pragma unroll
        for (int u = 0; u < U; u++)
        {
            const int t = ta + lane + 64 * u;
            colc[u] = Ai[t < tb ? t : 0];
            xc[u] = Ax[t < tb ? t : 0];
        }
        for (int j = 0; j < nq; j++)
        {
            const int a0 = ta, b0 = tb;
            const double v = __hiloint2double(__builtin_amdgcn_readlane(mvhi, j),
                                              __builtin_amdgcn_readlane(mvlo, j));
            const bool more = j + 1 < nq;
            const int jn = min(j + 1, 63);
            ta = more ? __builtin_amdgcn_readlane(ma, jn) : 0;
            tb = more ? __builtin_amdgcn_readlane(mb, jn) : 0;
            int coln[U];
            double xn[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int t = ta + lane + 64 * u;
                coln[u] = Ai[t < tb ? t : 0];
                xn[u] = Ax[t < tb ? t : 0];
            }
            // the U chunks' slots are distinct (one A row's columns): all
            // reads, then all writes (a slot without a term reads slot 0 and
            // writes nothing)
            unsigned o[U];
            double old[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                o[u] = a0 + lane + 64 * u < b0 ? slot(colc[u]) : (unsigned) WS;
                old[u] = a[o[u] < (unsigned) WS ? o[u] : 0u];
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (o[u] < (unsigned) WS) a[o[u]] = old[u] + v * xc[u];
            for (int t = a0 + lane + 64 * U; t < b0; t += 64)  // A rows longer than 64 U
            {
                const unsigned o = slot(Ai[t]);
                if (o < (unsigned) WS) a[o] = a[o] + v * Ax[t];
            }
            // (no code motion of memory operations across consecutive k)
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                colc[u] = coln[u];
                xc[u] = xn[u];
            }
        }
    }
    for (int i = lane; i < cnt; i += 64)
    {
        const unsigned o = SEARCH ? (unsigned) i : (unsigned) (Ci[cb + i] - lo);
        Cx[cb + i] = o < (unsigned) WS ? a[o] : 0.0;
    }
}

// Long rows with a dense accumulator of n doubles per workgroup (the
// reference's acc[n], gustavson.c:107-143): a persistent grid strides over the
// rows; per row the entries' slots are zeroed, L's row is walked in ascending
// k with a barrier per k (the columns of one A row are distinct, so its terms
// update distinct slots in parallel), then the row is gathered into C.
template <int G>
__global__ void __launch_bounds__(G)
    k_num_rowacc_dense(const int *__restrict__ rows, int nrows, const int *__restrict__ nrows_dev,
                       int Lm, const int *__restrict__ Lp,
                       const int *__restrict__ Li, const int *__restrict__ Lperm, int nL,
                       const int *__restrict__ Ap, const int *__restrict__ Ai,
                       const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
                       const double *__restrict__ Lval, const double *__restrict__ d,
                       const double *__restrict__ Ax, double *__restrict__ Cx,
                       double *__restrict__ scratch, int64_t n)
{
    double *acc = scratch + (int64_t) blockIdx.x * n;
    __shared__ int m_a[G], m_b[G];
    __shared__ double m_v[G];
    if (nrows_dev) nrows = *nrows_dev;
    for (int ri = blockIdx.x; ri < nrows; ri += gridDim.x)
    {
        const int row = rows[ri];
        const int64_t cbase = Cp[row], cnt = Cp[row + 1] - cbase;
        for (int64_t i = threadIdx.x; i < cnt; i += G) acc[Ci[cbase + i]] = 0.0;
        __syncthreads();
        RowCtx c = row_ctx(row, Lm, Lp, nL);
        // as k_num_rowacc: the chunk's A-row ranges and scale factors in LDS,
        // each A row's first G terms loaded before the previous row's updates
        for (int qb = c.q0; qb < c.q1; qb += G)
        {
            const int nqc = min(G, c.q1 - qb);
            if ((int) threadIdx.x < nqc)
            {
                const int q = qb + (int) threadIdx.x;
                const int k = Li[q] + c.koff;
                double v = Lval[Lperm ? Lperm[q] : q];
                if (d) v = v * d[k];
                m_a[threadIdx.x] = Ap[k];
                m_b[threadIdx.x] = Ap[k + 1];
                m_v[threadIdx.x] = v;
            }
            __syncthreads();
            int tc = m_a[0] + (int) threadIdx.x;
            bool hc = tc < m_b[0];
            int colc = Ai[hc ? tc : 0];
            double xc = Ax[hc ? tc : 0];
            for (int jq = 0; jq < nqc; jq++)
            {
                const int b = m_b[jq];
                const double v = m_v[jq];
                const bool more = jq + 1 < nqc;
                const int tn = (more ? m_a[jq + 1] : 0) + (int) threadIdx.x;
                const bool hn = more && tn < m_b[jq + 1];
                const int coln = Ai[hn ? tn : 0];
                const double xn = Ax[hn ? tn : 0];
                if (hc) acc[colc] = acc[colc] + v * xc;
                for (int t = tc + G; t < b; t += G)  // A rows longer than G
                {
                    const int j = Ai[t];
                    acc[j] = acc[j] + v * Ax[t];
                }
                __syncthreads();
                tc = tn;
                hc = hn;
                colc = coln;
                xc = xn;
            }
        }
        for (int64_t i = threadIdx.x; i < cnt; i += G) Cx[cbase + i] = acc[Ci[cbase + i]];
        __syncthreads();
    }
}

// Long rows by rank, for n up to RANK_MAX_WORDS * 32 columns and rows of up
// to RANK_MAX_ENTRIES entries: the row's output columns become an LDS column
// bitmap with per-word ranks, so a term's entry position is two LDS reads and
// a popcount.  Pass 1 marks the entries hit once and those hit more than once;
// pass 2 writes every once-hit entry directly (0 + its term, the fold of a
// single term) and lists the terms of the other entries as (entry, L index q,
// term).  The list is sorted by (entry, q) -- q ascends with k -- and each
// entry folds its terms in that order, so every entry is the same left fold
// in ascending k as the serial accumulators'.  Rows whose list outgrows its
// LDS capacity are appended to `spill` for the dense-accumulator kernel.  A
// persistent grid strides over the rows (one workgroup per CU: the bitmap and
// ranks take most of the LDS).
constexpr int RANK_MAX_WORDS = 16384;      // 524,288 columns
constexpr int RANK_MAX_ENTRIES = 65536;
constexpr int RANK_LIST = 1024;
constexpr int RU = 16;                     // terms per thread per step of the rank passes
constexpr int RU2 = 8;                     // the same in pass 2 (more registers per term)
template <bool B> struct BoolC
{
    static constexpr bool value = B;
};

__device__ __forceinline__ int rank_of(const unsigned *bm, const unsigned short *sub,
                                       const unsigned *sup, int col)
{
    const int w = col >> 5;
    const unsigned b = sup[w >> 10];
    const unsigned below = bm[w] & ((1u << (col & 31)) - 1u);
    return (int) (b + (unsigned short) (sub[w] - (unsigned short) b) + __popc(below));
}

// What the plan-time instance (PLAN = true) records, for rows of at most
// RANK_POS_MAX entries: every term's entry position within its row at
// rpos[roff[ri] + term index] (term index: the row's terms in ascending L
// index, then A entry), or, for a term of an entry of several terms,
// RANK_ITEM | its index in the row's item list -- the entry positions of
// those terms sorted by (entry, L index), at ipos[moff[ri] + i], i < mcnt[ri]
// <= RANK_LIST, in a region reserved by an atomic.  A longer row, or one whose
// list outgrows RANK_LIST, is marked in rskip and appended to spill.
constexpr int RANK_POS_MAX = 0x7FFF;
constexpr int RANK_ITEM = 0x8000;
struct RankRecord
{
    const int64_t *roff;
    uint16_t *rpos;
    unsigned char *rskip;
    int64_t *moff;
    int *mcnt;
    uint16_t *ipos;
    unsigned long long *mcount;
    int64_t mcap;
    int *err;
};

template <int G, bool PLAN = false>
__global__ void __launch_bounds__(G)
    k_num_rowacc_rank(const int *__restrict__ rows, int nrows, int Lm,
                      const int *__restrict__ Lp, const int *__restrict__ Li,
                      const int *__restrict__ Lperm, int nL, const int *__restrict__ Ap,
                      const int *__restrict__ Ai, const int64_t *__restrict__ Cp,
                      const int *__restrict__ Ci, const double *__restrict__ Lval,
                      const double *__restrict__ d, const double *__restrict__ Ax,
                      double *__restrict__ Cx, int nwords, int *__restrict__ spill,
                      int *__restrict__ nspill, RankRecord rec = RankRecord())
{
    extern __shared__ unsigned long long rank_lds[];
    unsigned long long *lkey = rank_lds;                       // RANK_LIST
    double *lval = (double *) (lkey + RANK_LIST);              // RANK_LIST
    // PLAN: the payload of a listed term, its index in rpos
    unsigned long long *lpay = (unsigned long long *) lval;
    unsigned *seen1 = (unsigned *) (lval + RANK_LIST);         // RANK_MAX_ENTRIES / 32
    unsigned *seen2 = seen1 + RANK_MAX_ENTRIES / 32;           // RANK_MAX_ENTRIES / 32
    unsigned *bm = seen2 + RANK_MAX_ENTRIES / 32;              // nwords
    const int nsup = (nwords + 1023) >> 10;
    unsigned *sup = bm + nwords;                               // nsup
    unsigned short *sub = (unsigned short *) (sup + nsup);     // nwords
    __shared__ Expand<G> ex;
    __shared__ int wsum[G / 64];
    __shared__ int nlist;
    __shared__ unsigned long long mbase;
    const int per = (nwords + G - 1) / G;
    for (int ri = blockIdx.x; ri < nrows; ri += gridDim.x)
    {
        const int row = rows[ri];
        const RowCtx c = row_ctx(row, Lm, Lp, nL);   // loads issued with Cp's
        const int64_t cbase = Cp[row];
        const int cnt = (int) (Cp[row + 1] - cbase);
        if (PLAN && cnt > RANK_POS_MAX)
        {
            // positions take 15 bits: a longer row spills
            if (threadIdx.x == 0)
            {
                spill[atomicAdd(nspill, 1)] = row;
                rec.rskip[ri] = 1;
            }
            continue;
        }
        const int swords = (cnt + 31) >> 5;
        for (int i = threadIdx.x; i < nwords; i += G) bm[i] = 0u;
        for (int i = threadIdx.x; i < swords; i += G)
        {
            seen1[i] = 0u;
            seen2[i] = 0u;
        }
        if (threadIdx.x == 0) nlist = 0;
        __syncthreads();
        for (int i0 = threadIdx.x; i0 < cnt; i0 += RU * G)
        {
            int col[RU];
#pragma unroll
            for (int u = 0; u < RU; u++)
            {
                const int i = i0 + u * G;
                const int cc = Ci[cbase + (i < cnt ? i : cnt - 1)];
                col[u] = i < cnt ? cc : -1;
            }
#pragma unroll
            for (int u = 0; u < RU; u++)
                if (col[u] >= 0) atomicOr(&bm[col[u] >> 5], 1u << (col[u] & 31));
        }
        __syncthreads();
        // ranks: a contiguous word range per thread, exclusive scan over threads
        {
            const int w0 = min(nwords, (int) threadIdx.x * per), w1 = min(nwords, w0 + per);
            int local = 0;
            for (int w = w0; w < w1; w++) local += __popc(bm[w]);
            int tot;
            int run = block_excl_scan_int<G>(local, &tot, wsum);
            for (int w = w0; w < w1; w++)
            {
                if ((w & 1023) == 0) sup[w >> 10] = (unsigned) run;
                sub[w] = (unsigned short) run;
                run += __popc(bm[w]);
            }
        }
        __syncthreads();
        // a row of one L chunk keeps pass 1's expansion for pass 2
        const bool one_chunk = c.q1 - c.q0 <= G;
        int tot1 = 0;
        // pass 1: entries hit once / more than once (RU terms per thread at a
        // time, their loads issued together: slots past the chunk repeat its
        // last term and are dropped, so no load sits behind a branch)
        for (int qb = c.q0; qb < c.q1; qb += G)
        {
            const int tot = expand_chunk<G>(ex, qb, c.q1, c.koff, Li, Ap);
            tot1 = tot;
            const int nq = min(G, c.q1 - qb);
            for (int f0 = threadIdx.x; f0 < tot; f0 += RU * G)
            {
                int col[RU];
#pragma unroll
                for (int u = 0; u < RU; u++)
                {
                    const int fi = f0 + u * G;
                    int t = 0, q, k;
                    expand_term<G>(ex, nq, fi < tot ? fi : tot - 1, t, q, k);
                    const int a = Ai[t];
                    col[u] = fi < tot ? a : -1;
                }
#pragma unroll
                for (int u = 0; u < RU; u++)
                {
                    if (col[u] < 0) continue;
                    const int pos = rank_of(bm, sub, sup, col[u]);
                    const unsigned bit = 1u << (pos & 31);
                    if (atomicOr(&seen1[pos >> 5], bit) & bit) atomicOr(&seen2[pos >> 5], bit);
                }
            }
            __syncthreads();
        }
        // pass 2: single-term entries written, the others' terms listed
        // (PLAN: every term's position recorded instead)
        int64_t tbase = PLAN ? rec.roff[ri] : 0;
        for (int qb = c.q0; qb < c.q1; qb += G)
        {
            const int tot = one_chunk ? tot1 : expand_chunk<G>(ex, qb, c.q1, c.koff, Li, Ap);
            const int nq = min(G, c.q1 - qb);
            if constexpr (PLAN)
            {
                for (int f0 = threadIdx.x; f0 < tot; f0 += RU2 * G)
                {
                    int col[RU2], qq[RU2];
#pragma unroll
                    for (int u = 0; u < RU2; u++)
                    {
                        const int fi = f0 + u * G;
                        int t = 0, q = 0, k = 0;
                        expand_term<G>(ex, nq, fi < tot ? fi : tot - 1, t, q, k);
                        const int a = Ai[t];
                        col[u] = fi < tot ? a : -1;
                        qq[u] = q;
                    }
#pragma unroll
                    for (int u = 0; u < RU2; u++)
                    {
                        if (col[u] < 0) continue;
                        const int pos = rank_of(bm, sub, sup, col[u]);
                        const bool single = !((seen2[pos >> 5] >> (pos & 31)) & 1u);
                        const int64_t at = tbase + f0 + u * G;
                        if (single) rec.rpos[at] = (uint16_t) pos;
                        else
                        {
                            // the item index is written once the list is sorted
                            const int i = atomicAdd(&nlist, 1);
                            if (i < RANK_LIST)
                            {
                                lkey[i] = ((unsigned long long) (unsigned) pos << 32) |
                                          (unsigned) qq[u];
                                lpay[i] = (unsigned long long) at;
                            }
                        }
                    }
                }
                tbase += tot;
                __syncthreads();
                continue;
            }
            for (int f0 = threadIdx.x; f0 < tot; f0 += RU2 * G)
            {
                int col[RU2], qq[RU2];
                double x[RU2];
                // unconditional loads (slots past the chunk repeat its last
                // term); the permutation / scaling choice is made outside the
                // unrolled loop so no branch separates the terms' loads
                auto load_terms = [&](auto perm_c, auto d_c) {
                    constexpr bool HP = decltype(perm_c)::value, HD = decltype(d_c)::value;
#pragma unroll
                    for (int u = 0; u < RU2; u++)
                    {
                        const int fi = f0 + u * G;
                        int t = 0, q = 0, k = 0;
                        expand_term<G>(ex, nq, fi < tot ? fi : tot - 1, t, q, k);
                        const int a = Ai[t];
                        double v = Lval[HP ? Lperm[q] : q];
                        if (HD) v = v * d[k];
                        const double xv = v * Ax[t];
                        col[u] = fi < tot ? a : -1;
                        qq[u] = q;
                        x[u] = xv;
                    }
                };
                if (Lperm)
                {
                    if (d) load_terms(BoolC<true>{}, BoolC<true>{});
                    else load_terms(BoolC<true>{}, BoolC<false>{});
                }
                else
                {
                    if (d) load_terms(BoolC<false>{}, BoolC<true>{});
                    else load_terms(BoolC<false>{}, BoolC<false>{});
                }
#pragma unroll
                for (int u = 0; u < RU2; u++)
                {
                    if (col[u] < 0) continue;
                    const int pos = rank_of(bm, sub, sup, col[u]);
                    if (!((seen2[pos >> 5] >> (pos & 31)) & 1u))
                    {
                        Cx[cbase + pos] = 0.0 + x[u];
                    }
                    else
                    {
                        const int i = atomicAdd(&nlist, 1);
                        if (i < RANK_LIST)
                        {
                            lkey[i] = ((unsigned long long) (unsigned) pos << 32) | (unsigned) qq[u];
                            lval[i] = x[u];
                        }
                    }
                }
            }
            __syncthreads();
        }
        const int nl = nlist;
        if (nl > RANK_LIST)
        {
            if (threadIdx.x == 0)
            {
                spill[atomicAdd(nspill, 1)] = row;
                if (PLAN) rec.rskip[ri] = 1;
            }
            __syncthreads();
            continue;
        }
        // sort the listed terms by (entry, q): bitonic over the next power of two
        int np = 1;
        while (np < nl) np <<= 1;
        for (int i = nl + (int) threadIdx.x; i < np; i += G)
        {
            lkey[i] = ~0ull;
            lpay[i] = 0ull;
        }
        __syncthreads();
        for (int kk = 2; kk <= np; kk <<= 1)
        {
            for (int j = kk >> 1; j > 0; j >>= 1)
            {
                for (int p = (int) threadIdx.x; p < np / 2; p += G)
                {
                    const int i = (p / j) * 2 * j + (p % j);
                    const int ixj = i + j;
                    const unsigned long long a = lkey[i], b = lkey[ixj];
                    const bool up = (i & kk) == 0;
                    if ((a > b) == up)
                    {
                        lkey[i] = b;
                        lkey[ixj] = a;
                        const unsigned long long va = lpay[i];
                        lpay[i] = lpay[ixj];
                        lpay[ixj] = va;
                    }
                }
                __syncthreads();
            }
        }
        if constexpr (PLAN)
        {
            // the sorted list becomes this row's region of the item list;
            // each listed term's rpos slot gets its item index
            if (threadIdx.x == 0) mbase = atomicAdd(rec.mcount, (unsigned long long) nl);
            __syncthreads();
            const unsigned long long mb = mbase;
            if (mb + (unsigned long long) nl > (unsigned long long) rec.mcap)
            {
                if (threadIdx.x == 0) atomicOr(rec.err, 1);
            }
            else
            {
                if (threadIdx.x == 0)
                {
                    rec.moff[ri] = (int64_t) mb;
                    rec.mcnt[ri] = nl;
                }
                for (int i = threadIdx.x; i < nl; i += G)
                {
                    rec.ipos[mb + i] = (uint16_t) (lkey[i] >> 32);
                    rec.rpos[(int64_t) lpay[i]] = (uint16_t) (RANK_ITEM | i);
                }
            }
            __syncthreads();
            continue;
        }
        // fold each listed entry's terms in ascending q
        for (int i = threadIdx.x; i < nl; i += G)
        {
            const unsigned pos = (unsigned) (lkey[i] >> 32);
            if (i > 0 && (unsigned) (lkey[i - 1] >> 32) == pos) continue;
            double acc = 0.0;
            for (int j = i; j < nl && (unsigned) (lkey[j] >> 32) == pos; j++) acc = acc + lval[j];
            Cx[cbase + pos] = acc;
        }
        __syncthreads();
    }
}

__global__ void k_rank_rows(const int *__restrict__ rows, int n,
                            const unsigned char *__restrict__ rskip,
                            const int64_t *__restrict__ roff, const int64_t *__restrict__ moff,
                            const int *__restrict__ mcnt, int Lm, const int *__restrict__ Lp,
                            int nL, const int64_t *__restrict__ Cp, RankRow *__restrict__ out)
{
    const int i = (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int row = rows[i];
    const RowCtx c = row_ctx(row, Lm, Lp, nL);
    RankRow r;
    r.cbase = Cp[row];
    r.tbase = roff[i];
    r.skip = rskip[i];
    r.mb = r.skip ? 0 : moff[i];
    r.nl = r.skip ? 0 : mcnt[i];
    r.q0 = c.q0;
    r.q1 = c.q1;
    r.koff = c.koff;
    r.cnt = (int) (Cp[row + 1] - r.cbase);
    out[i] = r;
}

// Refill of the rank bin from recorded positions.  A persistent grid, one
// workgroup per CU (the staging window takes most of the LDS), strides over
// the rows; a row is staged in LDS in windows of S entries (one window for
// rows of at most S entries): its terms in the recorded order (ascending L
// index, then A entry) -- 2-byte positions and A's values by row segment,
// both coalesced -- each single-term entry staged as 0 + (L value * d_k) * A
// value, each term of an entry of several terms kept at its item index; the
// items then fold in their sorted order (ascending L index per entry, so
// ((0 + x_1) + x_2) + ...), the accumulators' arithmetic; the window is
// written out with whole-line stores.  The next row's descriptor and its
// first chunk's L entries are loaded while the current row runs.
template <int G, int CH, int U>
__global__ void __launch_bounds__(G)
    k_num_rank_pos(const RankRow *__restrict__ rr, int nrows, const uint16_t *__restrict__ rpos,
                   const uint16_t *__restrict__ ipos, const int *__restrict__ Li,
                   const int *__restrict__ Lperm, const int *__restrict__ Ap,
                   const int2 *__restrict__ qrange, const double *__restrict__ Lval,
                   const double *__restrict__ d, const double *__restrict__ Ax,
                   double *__restrict__ Cx, int S)
{
    static_assert(CH <= G && CH % 64 == 0, "chunk of at most G L entries");
    extern __shared__ double stage[];  // S
    __shared__ Expand<CH> ex;          // a chunk of CH L entries, the first CH threads'
    __shared__ int wsum[G / 64];
    __shared__ double vq[CH];
    __shared__ double litem[RANK_LIST];
    __shared__ unsigned short lipos[RANK_LIST];
    int ri = blockIdx.x;
    if (ri >= nrows) return;
    RankRow cur = rr[ri];
    // the first chunk's L entry of this thread, for the current row: its A
    // row (k), that row's start and length, and L value * d_k
    // (qrange, when given: the A-row range of every L entry, recorded by the
    // plan, read in L order instead of gathered from Ap)
    auto slot_k = [&](const RankRow &r, int &kq, int &lq, int2 &qr) {
        const int q = r.q0 + (int) threadIdx.x;
        kq = -1;
        lq = 0;
        qr = int2{0, 0};
        if ((int) threadIdx.x < CH && q < r.q1)
        {
            kq = Li[q] + r.koff;
            lq = Lperm ? Lperm[q] : q;
            if (qrange) qr = qrange[q];
        }
    };
    auto slot_rest = [&](int kq, int lq, int2 qr, int &st, int &len, double &v) {
        st = 0;
        len = 0;
        v = 0.0;
        if (kq >= 0)
        {
            if (qrange)
            {
                st = qr.x;
                len = qr.y;
            }
            else
            {
                st = Ap[kq];
                len = Ap[kq + 1] - st;
            }
            v = Lval[lq];
            if (d) v = v * d[kq];
        }
    };
    int pk, pl, pst, plen;
    int2 pqr;
    double pv;
    slot_k(cur, pk, pl, pqr);
    slot_rest(pk, pl, pqr, pst, plen, pv);
    for (; ri < nrows; ri += gridDim.x)
    {
        const int rn = ri + gridDim.x;
        RankRow nxt = cur;
        if (rn < nrows) nxt = rr[rn];
        int nk = -1, nlq = 0;
        int2 nqr = int2{0, 0};
        bool pre_next = false;
        if (!cur.skip)
        {
            for (int i = threadIdx.x; i < cur.nl; i += G) lipos[i] = ipos[cur.mb + i];
            for (int w0 = 0; w0 < cur.cnt; w0 += S)
            {
                const int w1 = min(cur.cnt, w0 + S);
                int64_t tbase = cur.tbase;
                for (int qb = cur.q0; qb < cur.q1; qb += CH)
                {
                    int k = 0, st = 0, len = 0;
                    double v = 0.0;
                    if (w0 == 0 && qb == cur.q0)
                    {
                        k = pk < 0 ? 0 : pk;
                        st = pst;
                        len = plen;
                        v = pv;
                    }
                    else
                    {
                        const int q = qb + (int) threadIdx.x;
                        int kq = -1, lq = 0;
                        int2 qr = int2{0, 0};
                        if ((int) threadIdx.x < CH && q < cur.q1)
                        {
                            kq = Li[q] + cur.koff;
                            lq = Lperm ? Lperm[q] : q;
                            if (qrange) qr = qrange[q];
                        }
                        slot_rest(kq, lq, qr, st, len, v);
                        k = kq < 0 ? 0 : kq;
                    }
                    int tot;
                    const int ex_off = block_excl_scan_int<G>(len, &tot, wsum);
                    if ((int) threadIdx.x < CH)
                    {
                        ex.start[threadIdx.x] = st;
                        ex.k[threadIdx.x] = k;
                        ex.q[threadIdx.x] = (int) threadIdx.x;  // the slot: vq's index
                        ex.off[threadIdx.x] = ex_off;
                        vq[threadIdx.x] = v;
                    }
                    if (threadIdx.x == 0) ex.off[CH] = tot;
                    __syncthreads();
                    if (!pre_next && rn < nrows)
                    {
                        // first level of the next row's first chunk
                        slot_k(nxt, nk, nlq, nqr);
                        pre_next = true;
                    }
                    const int nq = min(CH, cur.q1 - qb);
                    for (int f0 = threadIdx.x; f0 < tot; f0 += U * G)
                    {
                        int p[U];
                        double x[U];
#pragma unroll
                        for (int u = 0; u < U; u++)
                        {
                            const int fi = f0 + u * G;
                            const int fc = fi < tot ? fi : tot - 1;
                            int t = 0, slot = 0, kk = 0;
                            expand_term<CH>(ex, nq, fc, t, slot, kk);
                            // positions are read once: kept out of L2's way
                            // of the re-read A values
                            const int pr = __builtin_nontemporal_load(rpos + tbase + fc);
                            x[u] = vq[slot] * Ax[t];
                            p[u] = fi < tot ? pr : -1;
                        }
#pragma unroll
                        for (int u = 0; u < U; u++)
                        {
                            const int pr = p[u];
                            if (pr < 0) continue;
                            if (pr & RANK_ITEM)
                            {
                                if (w0 == 0) litem[pr & (RANK_ITEM - 1)] = x[u];
                            }
                            else if (pr >= w0 && pr < w1) stage[pr - w0] = 0.0 + x[u];
                        }
                    }
                    tbase += tot;
                    __syncthreads();
                }
                // the entries of several terms in this window: the first item
                // of each folds the entry's items in order
                for (int i = threadIdx.x; i < cur.nl; i += G)
                {
                    const int pos = lipos[i];
                    if (pos < w0 || pos >= w1) continue;
                    if (i > 0 && lipos[i - 1] == pos) continue;
                    double acc = 0.0;
                    for (int j = i; j < cur.nl && lipos[j] == pos; j++) acc = acc + litem[j];
                    stage[pos - w0] = acc;
                }
                __syncthreads();
                // written out in 16-byte stores from the first 16-byte
                // boundary (an entry before it is stored alone)
                double *out = Cx + cur.cbase + w0;
                const int wn = w1 - w0;
                const int h = (int) (((uintptr_t) out >> 3) & 1);
                if (h && threadIdx.x == 0) __builtin_nontemporal_store(stage[0], out);
                const int npair = (wn - h) >> 1;
                dvec2 *out2 = (dvec2 *) (out + h);
                for (int j = threadIdx.x; j < npair; j += G)
                {
                    dvec2 v2;
                    v2.x = stage[h + 2 * j];
                    v2.y = stage[h + 2 * j + 1];
                    __builtin_nontemporal_store(v2, out2 + j);
                }
                if (((wn - h) & 1) && threadIdx.x == 0)
                    __builtin_nontemporal_store(stage[wn - 1], out + wn - 1);
                __syncthreads();
            }
        }
        if (rn >= nrows) break;
        if (!pre_next) slot_k(nxt, nk, nlq, nqr);
        slot_rest(nk, nlq, nqr, pst, plen, pv);
        pk = nk;
        cur = nxt;
    }
}

// Terms per row of the rank bin (for its offsets), its entry total and its
// longest row.
__global__ void k_rank_bin_sums(const int *__restrict__ rows, int n,
                                const int64_t *__restrict__ rterms, const int *__restrict__ rowcnt,
                                int64_t *__restrict__ out, unsigned long long *__restrict__ nnz,
                                unsigned *__restrict__ longest)
{
    const int i = (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    out[i] = rterms[rows[i]];
    atomicAdd(nnz, (unsigned long long) rowcnt[rows[i]]);
    atomicMax(longest, (unsigned) rowcnt[rows[i]]);
}

// A-row range (start, length) of every L entry of a one-block L.
__global__ void k_qrange(int64_t n, const int *__restrict__ Li, const int *__restrict__ Ap,
                         int2 *__restrict__ out)
{
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int k = Li[q];
    const int st = Ap[k];
    out[q] = int2{st, Ap[k + 1] - st};
}

__global__ void k_zero_int(int *p, int n)
{
    const int i = (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}

size_t rank_lds_bytes(int nwords)
{
    const int nsup = (nwords + 1023) >> 10;
    return (size_t) RANK_LIST * 16 + (size_t) (RANK_MAX_ENTRIES / 32) * 8 + (size_t) nwords * 4 +
           (size_t) nsup * 4 + (size_t) nwords * 2;
}

// ------------------------------------------------------------- host ----
struct Bins
{
    std::vector<int> count = std::vector<int>(NCLASS, 0);
    std::vector<int> off = std::vector<int>(NCLASS + 1, 0);
    DBuf<int> rows;
};

template <typename T>
void bin_rows(Bins &B, const T *cap, int64_t m, int nwords, const SymTune &tune, int mode = 0,
              int t1 = 0, int t2 = 0, const int *span = nullptr,
              const int64_t *terms = nullptr)
{
    DBuf<unsigned char> cls(m);
    DBuf<int> counts(NCLASS);
    dzero(counts.p, counts.bytes());
    unsigned grid = (unsigned) ((m + 255) / 256);
    if (m > 0)
        hipLaunchKernelGGL(k_classify<T>, dim3(grid), dim3(256), 0, stream(), m, cap, mode,
                           nwords, lds_bitmap_words(), tune.hash_max, tune.dens, t1, t2, span,
                           terms, cls.p, counts.p);
    SDL_HIP(hipGetLastError());
    d2h(B.count.data(), counts.p, NCLASS * sizeof(int));
    B.off[0] = 0;
    for (int c = 0; c < NCLASS; c++) B.off[c + 1] = B.off[c] + (c == 0 ? 0 : B.count[c]);
    B.rows.alloc(std::max(1, B.off[NCLASS]));
    std::vector<int> cur(NCLASS);
    for (int c = 0; c < NCLASS; c++) cur[c] = B.off[c];
    h2d(counts.p, cur.data(), NCLASS * sizeof(int));
    if (m > 0)
        hipLaunchKernelGGL(k_scatter_rows, dim3(grid), dim3(256), 0, stream(), m, cls.p,
                           counts.p, B.rows.p);
    SDL_HIP(hipGetLastError());
    sync();
}

template <int PASS>
void run_symbolic_class(int c, const Bins &B, const ProductPlan &P, int nwords, int *rowcnt,
                        int64_t *Cp, int *Ci)
{
    int nrows = B.count[c];
    if (nrows == 0) return;
    const int *rows = B.rows.p + B.off[c];
    const int nL = P.A.Am / P.L.blocks;
    const LView &L = P.L;
    const AView &A = P.A;
#define SDL_HASH_CASE(CLS, G, S)                                                          \
    case CLS:                                                                             \
        hipLaunchKernelGGL((k_sym_hash<G, S, PASS>), dim3(nrows), dim3(G), 0, stream(), rows, \
                           L.Lm, L.Lp, L.Li, nL, A.Ap, A.Ai, rowcnt, Cp, Ci,             \
                           g_sym_groups ? std::min(P.sym_sg, G) : 0);                    \
        break;
    switch (c)
    {
        SDL_HASH_CASE(1, 64, 64)
        SDL_HASH_CASE(2, 64, 128)
        SDL_HASH_CASE(3, 64, 256)
        SDL_HASH_CASE(4, 128, 512)
        SDL_HASH_CASE(5, 128, 1024)
        SDL_HASH_CASE(6, 256, 2048)
        SDL_HASH_CASE(7, 256, 4096)
        SDL_HASH_CASE(8, 256, 8192)
    case CLS_BITMAP:
    {
        // 1024 threads per row where the bitmap leaves room for their static
        // arrays (more waves to hide the term loads), else 256
        size_t lds = (size_t) nwords * 4;
        const bool wide = lds + 20 * 1024 <= 160 * 1024;
        auto fn = wide ? k_sym_bitmap<1024, PASS, false> : k_sym_bitmap<256, PASS, false>;
        SDL_HIP(hipFuncSetAttribute((const void *) fn,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
        hipLaunchKernelGGL(fn, dim3(nrows), dim3(wide ? 1024 : 256), lds, stream(), rows, nrows,
                           nwords, L.Lm, L.Lp, L.Li, nL, A.Ap, A.Ai, rowcnt, Cp, Ci,
                           (unsigned *) nullptr, g_sym_groups ? P.sym_sg : 0);
        break;
    }
    case CLS_WIN0 + 0:
    case CLS_WIN0 + 1:
    case CLS_WIN0 + 2:
    case CLS_WIN0 + 3:
    case CLS_WIN0 + 4:
    case CLS_WIN0 + 5:
    {
        SDL_REQUIRE(P.win_lo != nullptr, "symbolic: span window class without spans");
        const int sg = P.sym_sg;
        SDL_REQUIRE(sg >= 4 && sg <= 64 && (sg & (sg - 1)) == 0, "symbolic: window lanes per entry");
#define SDL_WIN_G(G, W)                                                                       \
    hipLaunchKernelGGL((k_sym_window<G, W, PASS>), dim3(nrows), dim3(G), 0, stream(), rows,   \
                       L.Lm, L.Lp, L.Li, nL, A.Ap, A.Ai, P.win_lo, sg, rowcnt, Cp, Ci)
        // one wavefront per row up to 1024-column windows (measured fastest on
        // banded and strided products: more rows in flight per CU), more for
        // the larger windows, whose rows carry more terms
        SDL_REQUIRE(sg <= 64, "symbolic: window lanes per entry");
        switch (c - CLS_WIN0)
        {
        case 0: SDL_WIN_G(64, 256); break;
        case 1: SDL_WIN_G(64, 512); break;
        case 2: SDL_WIN_G(64, 1024); break;
        case 3: SDL_WIN_G(128, 2048); break;
        case 4: SDL_WIN_G(256, 4096); break;
        case 5: SDL_WIN_G(256, 8192); break;
        default: break;
        }
#undef SDL_WIN_G
#undef SDL_WIN_CASE
        break;
    }
    case CLS_GLOBAL:
    {
        int grid = std::min(nrows, 512);
        DBuf<unsigned> scratch((size_t) grid * nwords);
        hipLaunchKernelGGL((k_sym_bitmap<256, PASS, true>), dim3(grid), dim3(256), 0,
                           stream(), rows, nrows, nwords, L.Lm, L.Lp, L.Li, nL, A.Ap, A.Ai,
                           rowcnt, Cp, Ci, scratch.p, 0);
        SDL_HIP(hipGetLastError());
        sync();
        break;
    }
    default: break;
    }
#undef SDL_HASH_CASE
    SDL_HIP(hipGetLastError());
}

// Count pass: each row's accumulator sized from cap[row] (rows with cap 0 are
// skipped); rowcnt[row] = the row's distinct columns.
void symbolic_count(ProductPlan &P, const int64_t *cap, int *rowcnt, int nwords, bool first)
{
    Bins B;
    bin_rows<int64_t>(B, cap, P.m_out, nwords, P.tune, 0, 0, 0,
                      g_sym_window ? P.win_span : nullptr, P.win_terms);
    for (int c = 1; c < NCLASS; c++) run_symbolic_class<0>(c, B, P, nwords, rowcnt, nullptr, nullptr);
    if (first)
    {
        P.stats.count_hash_rows = 0;
        for (int c = 1; c < CLS_BITMAP; c++) P.stats.count_hash_rows += B.count[c];
        P.stats.count_bitmap_rows = B.count[CLS_BITMAP];
        P.stats.count_global_rows = B.count[CLS_GLOBAL];
        P.stats.count_window_rows = 0;
        for (int c = CLS_WIN0; c < NCLASS; c++) P.stats.count_window_rows += B.count[c];
    }
    sync();
}

// Scan of the exact counts, then the fill pass (tables sized from the exact
// count).
void symbolic_fill(ProductPlan &P, const int *rowcnt, int nwords)
{
    const int64_t m_out = P.m_out;
    P.nnz = scan_i32_to_i64(rowcnt, P.Cp.p, m_out, true);
    P.stats.nnz = P.nnz;
    P.Ci.alloc(P.nnz);
    Bins B2;
    bin_rows<int>(B2, rowcnt, m_out, nwords, P.tune, 0, 0, 0,
                  g_sym_window ? P.win_span : nullptr, P.win_terms);
    for (int c = 1; c < NCLASS; c++)
    {
        run_symbolic_class<1>(c, B2, P, nwords, nullptr, P.Cp.p, P.Ci.p);
        if (c >= 1 && c <= 8) P.stats.hash_rows += B2.count[c];
        if (c >= CLS_WIN0) P.stats.window_rows += B2.count[c];
    }
    P.stats.bitmap_rows = B2.count[CLS_BITMAP];
    P.stats.global_rows = B2.count[CLS_GLOBAL];
    sync();
}

// Count pass with accumulators sized from the term bound, scan, fill pass.
// Count pass sized from each row's exact nnz bound min(terms, span), then
// the scan and the fill pass.
void symbolic_exact(ProductPlan &P, const int64_t *rbound, int *rowcnt, int nwords)
{
    symbolic_count(P, rbound, rowcnt, nwords, true);
    symbolic_fill(P, rowcnt, nwords);
}

// Count pass with accumulators sized from the predicted bins (never above the
// term bound), a second count from the term bound for the rows whose count
// exceeded the predicted bound, then the scan and fill pass of the exact path.
// With no predictor every bin is unbounded and this is the exact path.
void symbolic_predicted(ProductPlan &P, const RowPredictor *pred, const int64_t *rterms,
                        const int64_t *rbound, int *rowcnt, int nwords)
{
    const int64_t m_out = P.m_out;
    const unsigned grid = (unsigned) ((m_out + 255) / 256);
    // rows a prediction could help: counted first, so that a product without
    // any takes the exact path with nothing allocated for the prediction
    int nc = 0;
    if (pred)
    {
        DBuf<int> ncand(1);
        dzero(ncand.p, sizeof(int));
        hipLaunchKernelGGL(k_pred_candidates, dim3(grid), dim3(256), 0, stream(), m_out, rterms,
                           rbound, nwords, lds_bitmap_words(), P.tune.hash_max, P.tune.dens,
                           (int *) nullptr, ncand.p);
        SDL_HIP(hipGetLastError());
        d2h(&nc, ncand.p, sizeof(int));
    }
    P.stats.candidate_rows = nc;
    if (nc == 0)
    {
        // no candidate: the exact path itself; every non-empty row kept its
        // term bound (the count pass's classified rows)
        symbolic_exact(P, rbound, rowcnt, nwords);
        P.stats.predicted_rows = 0;
        P.stats.overflow_rows = 0;
        P.stats.misbin_rows = 0;
        P.stats.exact_rows = P.stats.count_hash_rows + P.stats.count_bitmap_rows +
                             P.stats.count_global_rows + P.stats.count_window_rows;
        return;
    }
    DBuf<unsigned char> pbin(m_out);
    DBuf<int64_t> cap(m_out), recount(m_out);
    // every row starts from its term bound (the exact path's sizes); the
    // candidates get features, the network and a cost decision
    SDL_HIP(hipMemsetAsync(pbin.p, PBIN_NONE, (size_t) m_out, stream()));
    d2d(cap.p, rbound, (size_t) m_out * sizeof(int64_t));
    {
        DBuf<int> list(nc), ncand(1);
        dzero(ncand.p, sizeof(int));
        hipLaunchKernelGGL(k_pred_candidates, dim3(grid), dim3(256), 0, stream(), m_out, rterms,
                           rbound, nwords, lds_bitmap_words(), P.tune.hash_max, P.tune.dens,
                           list.p, ncand.p);
        SDL_HIP(hipGetLastError());
        {
            DBuf<float> F((size_t) nc * MLP_IN), Z((size_t) nc * MLP_OUT);
            hipLaunchKernelGGL(k_row_features, dim3((unsigned) ((nc + 3) / 4)), dim3(256), 0,
                               stream(), (int64_t) nc, list.p, P.L.Lm, P.L.Lp, P.L.Li,
                               P.A.Am / P.L.blocks, P.A.Ap, P.A.Ai, P.A.An, F.p);
            SDL_HIP(hipGetLastError());
            mlp_forward(*pred, F.p, nc, Z.p, stream());
            hipLaunchKernelGGL(k_pred_cap, dim3((unsigned) ((nc + 255) / 256)), dim3(256), 0,
                               stream(), (int64_t) nc, list.p, rterms, rbound, F.p, Z.p,
                               pred->tau, nwords, lds_bitmap_words(), P.tune.hash_max,
                               P.tune.dens, cap.p, pbin.p);
            SDL_HIP(hipGetLastError());
        }
    }
    DBuf<unsigned long long> st(4);
    {
        symbolic_count(P, cap.p, rowcnt, nwords, true);
        DBuf<unsigned long long> nrec(1);
        dzero(nrec.p, nrec.bytes());
        hipLaunchKernelGGL(k_pred_recount, dim3(grid), dim3(256), 0, stream(), m_out, rowcnt,
                           cap.p, rbound, recount.p, nrec.p);
        SDL_HIP(hipGetLastError());
        unsigned long long nr = 0;
        d2h(&nr, nrec.p, sizeof nr);
        // the second count only when a lowered bound was exceeded
        if (nr > 0) symbolic_count(P, recount.p, rowcnt, nwords, false);
        symbolic_fill(P, rowcnt, nwords);
    }
    dzero(st.p, st.bytes());
    hipLaunchKernelGGL(k_pred_stats, dim3(grid), dim3(256), 0, stream(), m_out, cap.p, recount.p,
                       pbin.p, rbound, P.Cp.p, st.p);
    SDL_HIP(hipGetLastError());
    unsigned long long h[4];
    d2h(h, st.p, sizeof(h));
    P.stats.predicted_rows = (long long) h[0];
    P.stats.overflow_rows = (long long) h[1];
    P.stats.misbin_rows = pred ? (long long) h[2] : 0;
    P.stats.exact_rows = (long long) h[3];
}

} // namespace

int64_t product_count_terms(const LView &L, const AView &A, int64_t *row_terms_out)
{
    int64_t m_out = (int64_t) L.Lm * L.blocks;
    if (m_out == 0) return 0;
    DBuf<int64_t> tmp;
    int64_t *terms = row_terms_out;
    if (!terms)
    {
        tmp.alloc(m_out);
        terms = tmp.p;
    }
    hipLaunchKernelGGL(k_row_terms, dim3((unsigned) ((m_out + 15) / 16)), dim3(256), 0,
                       stream(), m_out, L.Lm, L.Lp, L.Li, A.Am / L.blocks, A.Ap,
                       (const int2 *) nullptr, terms, (int64_t *) nullptr);
    SDL_HIP(hipGetLastError());
    DBuf<int64_t> off(m_out + 1);
    return scan_i64(terms, off.p, m_out, true);
}

void product_row_features(const LView &L, const AView &A, float *F, hipStream_t s)
{
    const int64_t m_out = (int64_t) L.Lm * L.blocks;
    if (m_out == 0) return;
    hipLaunchKernelGGL(k_row_features, dim3((unsigned) ((m_out + 3) / 4)), dim3(256), 0, s,
                       m_out, (const int *) nullptr, L.Lm, L.Lp, L.Li, A.Am / L.blocks, A.Ap,
                       A.Ai, A.An, F);
    SDL_HIP(hipGetLastError());
}

// LDS of k_num_rank_pos<RP_G, RP_CH, RP_U> beside its staging window, and the
// window: as many entries as the rest of the CU's 160 KB LDS holds, or the
// bin's longest row when shorter.
constexpr int RP_G = 1024, RP_CH = 256, RP_U = 8;
constexpr int CU_LDS_BYTES = 160 * 1024;

// Records the rank bin's term positions (ProductPlan::rankpos) when the bin's
// terms are at most 1.25x its entries -- the entries of several terms, and so
// the items, stay few -- and the positions take at most 48 GB.  Otherwise the
// bin keeps the per-refill ranks.
void record_rank_positions(ProductPlan &P, const LView &L, const AView &A, const int64_t *rterms,
                           const int *rowcnt, int mid)
{
    const int *rows = P.rbins.p + P.rbin_off[1];
    DBuf<int64_t> rt(mid);
    DBuf<unsigned long long> nz(1);
    DBuf<unsigned> longest(1);
    dzero(nz.p, nz.bytes());
    dzero(longest.p, longest.bytes());
    hipLaunchKernelGGL(k_rank_bin_sums, dim3((unsigned) ((mid + 255) / 256)), dim3(256), 0,
                       stream(), rows, mid, rterms, rowcnt, rt.p, nz.p, longest.p);
    SDL_HIP(hipGetLastError());
    DBuf<int64_t> roff((size_t) mid + 1);
    const int64_t T = scan_i64(rt.p, roff.p, mid, true);
    unsigned long long N = 0;
    unsigned lmax = 0;
    d2h(&N, nz.p, sizeof N);
    d2h(&lmax, longest.p, sizeof lmax);
    const int64_t extra = T - (int64_t) N;
    if (T <= 0 || 4 * extra > (int64_t) N || 2 * T > ((int64_t) 48 << 30)) return;
    P.roff = std::move(roff);
    P.rpos.alloc((size_t) T);
    P.rskip.alloc(mid);
    P.moff.alloc(mid);
    P.mcnt.alloc(mid);
    dzero(P.rskip.p, P.rskip.bytes());
    dzero(P.mcnt.p, P.mcnt.bytes());
    // items <= 2 (terms - entries) (an entry of c >= 2 terms lists c of
    // them), and <= RANK_LIST per row
    const int64_t mcap =
        std::max<int64_t>(1, std::min<int64_t>(2 * extra, (int64_t) mid * RANK_LIST));
    P.ipos.alloc(mcap);
    DBuf<unsigned long long> mcount(1);
    DBuf<int> err(1);
    dzero(mcount.p, mcount.bytes());
    dzero(err.p, err.bytes());
    dzero(P.nspill.p, sizeof(int));
    RankRecord rec;
    rec.roff = P.roff.p;
    rec.rpos = P.rpos.p;
    rec.rskip = P.rskip.p;
    rec.moff = P.moff.p;
    rec.mcnt = P.mcnt.p;
    rec.ipos = P.ipos.p;
    rec.mcount = mcount.p;
    rec.mcap = mcap;
    rec.err = err.p;
    const int nwords = (int) ((P.n + 31) / 32);
    SDL_HIP(hipFuncSetAttribute((const void *) k_num_rowacc_rank<1024, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int) rank_lds_bytes(RANK_MAX_WORDS)));
    const int nL = A.Am / L.blocks;
    hipLaunchKernelGGL((k_num_rowacc_rank<1024, true>), dim3(P.rank_grid), dim3(1024),
                       rank_lds_bytes(nwords), stream(), rows, mid, L.Lm, L.Lp, L.Li, L.Lperm,
                       nL, A.Ap, A.Ai, P.Cp.p, P.Ci.p, (const double *) nullptr,
                       (const double *) nullptr, (const double *) nullptr, (double *) nullptr,
                       nwords, P.spill.p, P.nspill.p, rec);
    SDL_HIP(hipGetLastError());
    int herr = 0;
    unsigned long long nm = 0;
    d2h(&herr, err.p, sizeof herr);
    d2h(&nm, mcount.p, sizeof nm);
    SDL_REQUIRE(herr == 0 && (int64_t) nm <= mcap, "rank positions: item list overflow");
    P.nmulti = (int64_t) nm;
    if (L.blocks == 1)
    {
        // every L entry's A-row range (one block: an L entry always meets
        // the same A row)
        int nle = 0;
        d2h(&nle, L.Lp + L.Lm, sizeof nle);
        const int64_t nL_entries = nle;
        P.qrange.alloc((size_t) std::max<int64_t>(nL_entries, 1));
        if (nL_entries > 0)
            hipLaunchKernelGGL(k_qrange, dim3((unsigned) ((nL_entries + 255) / 256)), dim3(256),
                               0, stream(), nL_entries, L.Li, A.Ap, P.qrange.p);
        SDL_HIP(hipGetLastError());
    }
    P.rrows.alloc(mid);
    hipLaunchKernelGGL(k_rank_rows, dim3((unsigned) ((mid + 255) / 256)), dim3(256), 0, stream(),
                       rows, mid, P.rskip.p, P.roff.p, P.moff.p, P.mcnt.p, L.Lm, L.Lp, nL,
                       P.Cp.p, P.rrows.p);
    SDL_HIP(hipGetLastError());
    hipFuncAttributes fa;
    SDL_HIP(hipFuncGetAttributes(&fa, (const void *) k_num_rank_pos<RP_G, RP_CH, RP_U>));
    const int wmax = (CU_LDS_BYTES - (int) fa.sharedSizeBytes - 2048) / 8 / 256 * 256;
    P.rank_window = std::min(wmax, (int) std::min<unsigned>(lmax, RANK_POS_MAX) + 255 & ~255);
    P.rank_window = std::max(P.rank_window, 256);
    SDL_HIP(hipFuncSetAttribute((const void *) k_num_rank_pos<RP_G, RP_CH, RP_U>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                P.rank_window * (int) sizeof(double)));
    P.rankpos = true;
}

void product_plan(ProductPlan &P, const LView &L, const AView &A, SymMode mode,
                  const RowPredictor *pred, bool allow_pull, bool force_pull)
{
    SDL_REQUIRE(L.blocks >= 1 && A.Am % L.blocks == 0, "product: A rows not a multiple of blocks");
    P.L = L;
    P.A = A;
    P.m_out = (int64_t) L.Lm * L.blocks;
    P.n = A.An;
    const int64_t m_out = P.m_out;
    const int nL = A.Am / L.blocks;
    const int nwords = (A.An + 31) / 32;
    P.stats = SymStats();
    P.stats.rows = m_out;
    P.Cp.alloc(m_out + 1);
    if (m_out == 0)
    {
        dzero(P.Cp.p, sizeof(int64_t));
        P.nnz = 0;
        P.terms = 0;
        P.Ci.alloc(0);
        return;
    }

    // per-row term bound, column span and the term scan
    DBuf<int64_t> rterms(m_out);
    DBuf<int64_t> rbound(m_out);
    DBuf<int> wlo, wspan;
    if (g_sym_window)
    {
        wlo.alloc(m_out);
        wspan.alloc(m_out);
    }
    {
        DBuf<int2> arange(std::max(A.Am, 1));
        if (A.Am > 0)
            hipLaunchKernelGGL(k_arow_range, dim3((unsigned) ((A.Am + 3) / 4)), dim3(256), 0,
                               stream(), A.Am, A.Ap, A.Ai, arange.p);
        hipLaunchKernelGGL(k_row_terms, dim3((unsigned) ((m_out + 15) / 16)), dim3(256), 0,
                           stream(), m_out, L.Lm, L.Lp, L.Li, nL, A.Ap, arange.p, rterms.p,
                           rbound.p, wlo.p, wspan.p);
        SDL_HIP(hipGetLastError());
    }
    SDL_HIP(hipGetLastError());
    DBuf<int64_t> toff(m_out + 1);
    P.terms = scan_i64(rterms.p, toff.p, m_out, true);
    P.stats.terms = P.terms;
    {
        int nle = 0;
        d2h(&nle, L.Lp + L.Lm, sizeof(int));
        P.mean_arow = (double) P.terms / std::max(1.0, (double) nle * L.blocks);
    }
    {
        // lanes per L entry of the window and LDS-bitmap kernels: a power of two in [4, 64]
        // near 1/16 of the mean A-row length (terms per L entry): groups of a
        // few lanes keep many L entries in flight, and four unrolled slots per
        // lane still cover most of a row in a pass (measured on banded rows of
        // 17..129 entries)
        const double mean = P.mean_arow;
        int sg = 4;
        while (sg < 64 && 32.0 * sg <= mean) sg <<= 1;
        P.sym_sg = g_sym_window_lanes ? g_sym_window_lanes : sg;
    }
    if (g_sym_window)
    {
        P.win_lo = wlo.p;
        P.win_span = wspan.p;
        P.win_terms = rterms.p;
    }

    // thresholds: forced (the tuner), the tuning table's for this shape, or
    // the defaults
    {
        auto lg = [](double v) { return v >= 1.0 ? (int) std::floor(std::log2(v)) : 0; };
        P.sig.lg_rows = lg((double) m_out);
        P.sig.lg_cols = lg((double) A.An);
        P.sig.lg_tpr = lg((double) P.terms / (double) m_out);
        RefillCfg rc;
        SymTune t;
        if (g_force_tune) P.tune = *g_force_tune;
        else if (tune_lookup(P.sig, &t, &rc))
        {
            P.tune = t;
            if (rc.variant >= 0) P.tuned = rc;
        }
    }

    // structure: count pass sized from the term bound or the predicted bins, fill pass
    DBuf<int> rowcnt(m_out);
    dzero(rowcnt.p, rowcnt.bytes());
    if (mode == SymMode::Predicted)
        symbolic_predicted(P, pred, rterms.p, rbound.p, rowcnt.p, nwords);
    else symbolic_exact(P, rbound.p, rowcnt.p, nwords);
    P.win_lo = P.win_span = nullptr;
    P.win_terms = nullptr;

    // refill schedule (reference policy: a plan iff terms <= 3 nnz, gustavson.h:45-52)
    P.pull = allow_pull && P.terms > 0 && (force_pull || P.terms <= 3 * P.nnz) &&
             P.terms < (int64_t) 0x7fffffff && P.nnz < (int64_t) 0x7fffffff;
    if (P.pull)
    {
        P.ep.alloc(P.nnz + 1);
        P.tA.alloc(P.terms);
        P.tB.alloc(P.terms);
        P.tK.alloc(P.terms);
        if (L.blocks == 1) P.tq.alloc(P.terms);  // the tile refill's L indices
        // rows binned by nnz: LDS lists up to 2048 entries, global beyond
        DBuf<int> err(1);
        dzero(err.p, err.bytes());
        Bins BT;
        bin_rows<int>(BT, rowcnt.p, m_out, nwords, P.tune, 1, 128, 2048);
        DBuf<int> gscratch;
        for (int c = 1; c <= 3; c++)
        {
            int nrows = BT.count[c];
            if (!nrows) continue;
            const int *rows = BT.rows.p + BT.off[c];
            if (c == 1)
                hipLaunchKernelGGL((k_terms<64, 128>), dim3(nrows), dim3(64), 0, stream(), rows,
                                   L.Lm, L.Lp, L.Li, L.Lperm, nL, A.Ap, A.Ai, P.Cp.p, P.Ci.p,
                                   toff.p, P.ep.p, P.tA.p, P.tB.p, P.tK.p, P.tq.p, (int *) nullptr,
                                   err.p);
            else if (c == 2)
                hipLaunchKernelGGL((k_terms<256, 2048>), dim3(nrows), dim3(256), 0, stream(),
                                   rows, L.Lm, L.Lp, L.Li, L.Lperm, nL, A.Ap, A.Ai, P.Cp.p,
                                   P.Ci.p, toff.p, P.ep.p, P.tA.p, P.tB.p, P.tK.p, P.tq.p,
                                   (int *) nullptr, err.p);
            else
            {
                gscratch.alloc(P.nnz);
                hipLaunchKernelGGL((k_terms<256, 0>), dim3(nrows), dim3(256), 0, stream(), rows,
                                   L.Lm, L.Lp, L.Li, L.Lperm, nL, A.Ap, A.Ai, P.Cp.p, P.Ci.p,
                                   toff.p, P.ep.p, P.tA.p, P.tB.p, P.tK.p, P.tq.p, gscratch.p,
                                   err.p);
            }
            SDL_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_set_last, dim3(1), dim3(64), 0, stream(), P.ep.p, P.nnz, P.terms);
        SDL_HIP(hipGetLastError());
        int herr = 0;
        d2h(&herr, err.p, sizeof(int));
        SDL_REQUIRE(herr == 0, "product plan: a term fell outside its row's structure");
        if (P.tq.p)
        {
            const int ntiles = (int) ((P.nnz + 255) / 256);
            P.tqr.alloc(ntiles);
            hipLaunchKernelGGL(k_tile_qrange, dim3((unsigned) ((ntiles + 255) / 256)), dim3(256), 0,
                               stream(), ntiles, P.nnz, m_out, P.Cp.p, L.Lp, P.tqr.p);
            SDL_HIP(hipGetLastError());
            P.tile_refill = true;
        }
    }
    else
    {
        // row-accumulator bins: rows with <= rowacc_cut (<= 1024) entries
        // accumulate in LDS; rows up to rank_cut (<= RANK_MAX_ENTRIES) entries
        // by rank when the column bitmap fits LDS; longer rows in dense
        // accumulators
        Bins BR;
        bin_rows<int>(BR, rowcnt.p, m_out, nwords, P.tune, 1, P.tune.rowacc_cut,
                      std::min(P.tune.rank_cut, RANK_MAX_ENTRIES));
        const int small = BR.count[1], mid = BR.count[2], large = BR.count[3];
        P.rbins.alloc(std::max(1, small + mid + large));
        P.rbin_off[0] = 0;
        P.rbin_off[1] = small;
        P.rbin_off[2] = small + mid;
        P.rbin_off[3] = small + mid + large;
        d2d(P.rbins.p, BR.rows.p + BR.off[1], (size_t) small * sizeof(int));
        d2d(P.rbins.p + small, BR.rows.p + BR.off[2], (size_t) mid * sizeof(int));
        d2d(P.rbins.p + small + mid, BR.rows.p + BR.off[3], (size_t) large * sizeof(int));
        // the LDS-accumulator rows by column span: rows within 256 / 512 /
        // 1024 columns take the window accumulators (after the others)
        for (int c = 0; c < RW_CLASSES; c++) P.rwin_off[c] = c == 0 ? 0 : small;
        if (g_rowacc_window && small > 0)
        {
            DBuf<unsigned char> cls(small);
            DBuf<int> counts(RW_CLASSES), tmp(small);
            dzero(counts.p, counts.bytes());
            const unsigned g = (unsigned) ((small + 255) / 256);
            hipLaunchKernelGGL(k_rowacc_span_class, dim3(g), dim3(256), 0, stream(), small,
                               P.rbins.p, P.Cp.p, P.Ci.p, cls.p, counts.p);
            SDL_HIP(hipGetLastError());
            int cnt[RW_CLASSES];
            d2h(cnt, counts.p, sizeof(cnt));
            int cur[RW_CLASSES] = {0};
            for (int c = 1; c < RW_CLASSES; c++)
            {
                cur[c] = P.rwin_off[c - 1];
                P.rwin_off[c] = P.rwin_off[c - 1] + cnt[c];
            }
            SDL_REQUIRE(P.rwin_off[RW_CLASSES - 1] == small,
                        "product plan: window classes do not cover the rows");
            h2d(counts.p, cur, sizeof(cur));
            hipLaunchKernelGGL(k_scatter_rows, dim3(g), dim3(256), 0, stream(), (int64_t) small,
                               cls.p, counts.p, tmp.p, (const int *) P.rbins.p);
            SDL_HIP(hipGetLastError());
            d2d(P.rbins.p, tmp.p, (size_t) small * sizeof(int));
        }
        P.rank = mid > 0 && nwords <= RANK_MAX_WORDS;
        if (P.rank)
        {
            P.spill.alloc(mid);
            P.nspill.alloc(1);
            int dev = 0, cus = 0;
            SDL_HIP(hipGetDevice(&dev));
            SDL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            P.rank_grid = std::max(1, std::min(mid, cus));
        }
        // dense accumulators for the long rows (and rows spilled by the rank
        // kernel): at most 2 GB and 1024 workgroups, at least 64, else C itself
        // is the accumulator
        const int64_t budget = (int64_t) 2 << 30;
        const int64_t fit = budget / (8 * std::max<int64_t>(P.n, 1));
        const int dense_rows = large + mid;
        if (dense_rows > 0 && fit >= 64)
        {
            P.dense_grid = (int) std::min<int64_t>({(int64_t) dense_rows, fit, 1024});
            P.acc.alloc((size_t) P.dense_grid * std::max<int64_t>(P.n, 1));
        }
        if (P.rank && !P.rank_refill) record_rank_positions(P, L, A, rterms.p, rowcnt.p, mid);
        // rank-bin rows without recorded positions (many terms per entry)
        // whose columns span at most 4096: the window accumulator, one
        // wavefront per workgroup, first in the bin
        for (int c = 0; c < 4; c++) P.rwide_off[c] = 0;
        if (g_rowacc_window && mid > 0 && !P.rankpos && !P.rank_refill)
        {
            int *mrows = P.rbins.p + P.rbin_off[1];
            DBuf<unsigned char> cls(mid);
            DBuf<int> counts(5), tmp(mid);
            dzero(counts.p, counts.bytes());
            const unsigned g = (unsigned) ((mid + 255) / 256);
            hipLaunchKernelGGL(k_rowacc_wide_class, dim3(g), dim3(256), 0, stream(), mid, mrows,
                               P.Cp.p, P.Ci.p, cls.p, counts.p);
            SDL_HIP(hipGetLastError());
            int cnt[5];
            d2h(cnt, counts.p, sizeof(cnt));
            if (cnt[2] + cnt[3] + cnt[4] > 0)
            {
                // order: span <= 2048 (class 2), <= 4096 (class 3), <= 8192
                // (class 4), then the rest (class 1); cur[c] is class c's
                // first slot
                int cur[5] = {0, cnt[2] + cnt[3] + cnt[4], 0, cnt[2], cnt[2] + cnt[3]};
                P.rwide_off[1] = cnt[2];
                P.rwide_off[2] = cnt[2] + cnt[3];
                P.rwide_off[3] = cnt[2] + cnt[3] + cnt[4];
                h2d(counts.p, cur, sizeof(cur));
                hipLaunchKernelGGL(k_scatter_rows, dim3(g), dim3(256), 0, stream(), (int64_t) mid,
                                   cls.p, counts.p, tmp.p, (const int *) mrows);
                SDL_HIP(hipGetLastError());
                d2d(mrows, tmp.p, (size_t) mid * sizeof(int));
            }
        }
        sync();
    }
}

void product_numeric(const ProductPlan &P, const double *Lval, const double *d,
                     const double *Ax, double *Cx, hipStream_t s)
{
    if (P.nnz == 0) return;
    if (P.pull && P.tile_refill && P.tq.p)
    {
        hipLaunchKernelGGL(k_num_pull_tile<false>, dim3((unsigned) ((P.nnz + 255) / 256)),
                           dim3(256), 0, s, P.nnz, P.tqr.p, P.ep.p, P.tq.p, P.tA.p, P.L.Li,
                           P.L.Lperm, Lval, d, Ax, Cx);
    }
    else if (P.pull)
    {
        hipLaunchKernelGGL(k_num_pull, dim3((unsigned) ((P.nnz + 255) / 256)), dim3(256), 0, s,
                           P.nnz, P.ep.p, P.tA.p, P.tB.p, P.tK.p, Lval, d, Ax, Cx);
    }
    else
    {
        const int nL = P.A.Am / P.L.blocks;
        const int n0 = P.rbin_off[1] - P.rbin_off[0];
        const int *long_rows = P.rbins.p + P.rbin_off[1];
        int n1 = P.rbin_off[3] - P.rbin_off[1];
        const int nsearch = P.rwin_off[1];
        if (nsearch)
            hipLaunchKernelGGL((k_num_rowacc<256, 1024>), dim3(nsearch), dim3(256), 0, s,
                               P.rbins.p, P.L.Lm, P.L.Lp, P.L.Li, P.L.Lperm, nL, P.A.Ap, P.A.Ai,
                               P.Cp.p, P.Ci.p, Lval, d, Ax, Cx);
        for (int c = 2; c < RW_CLASSES && n0; c++)
        {
            const int nr = P.rwin_off[c] - P.rwin_off[c - 1];
            if (nr == 0) continue;
            const int *wr = P.rbins.p + P.rwin_off[c - 1];
            // A-row chunks of 64 loaded ahead: enough for the mean A row
            const int u = P.mean_arow <= 64 ? 1 : P.mean_arow <= 128 ? 2 : P.mean_arow <= 192 ? 3 : 4;
#define SDL_RW(WS, SR)                                                                      \
    (u == 1 ? k_num_rowacc_win<WS, 1, SR> : u == 2 ? k_num_rowacc_win<WS, 2, SR>            \
                                          : u == 3 ? k_num_rowacc_win<WS, 3, SR>            \
                                                   : k_num_rowacc_win<WS, 4, SR>)
            auto fn = c == 2 ? SDL_RW(256, false) : c == 3 ? SDL_RW(512, false)
                    : c == 4 ? SDL_RW(1024, false) : c == 5 ? SDL_RW(256, true)
                    : c == 6 ? SDL_RW(512, true) : SDL_RW(1024, true);
#undef SDL_RW
            hipLaunchKernelGGL(fn, dim3((unsigned) ((nr + 3) / 4)), dim3(256), 0, s, wr, nr,
                               P.L.Lm, P.L.Lp, P.L.Li, P.L.Lperm, nL, P.A.Ap, P.A.Ai, P.Cp.p,
                               P.Ci.p, Lval, d, Ax, Cx);
        }
        // rank-bin rows in the wide windows (first in the bin)
        const int nwide = P.rwide_off[3];
        static bool win8k_attr = false;
        if (!win8k_attr && P.rwide_off[3] > P.rwide_off[2])
        {
            for (auto f : {(const void *) k_num_rowacc_win<8192, 1, false, 1>,
                           (const void *) k_num_rowacc_win<8192, 2, false, 1>,
                           (const void *) k_num_rowacc_win<8192, 3, false, 1>,
                           (const void *) k_num_rowacc_win<8192, 4, false, 1>})
                SDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            8192 * (int) sizeof(double)));
            win8k_attr = true;
        }
        for (int c = 1; c <= 3 && nwide; c++)
        {
            const int nr = P.rwide_off[c] - P.rwide_off[c - 1];
            if (nr == 0) continue;
            const int *wr = long_rows + P.rwide_off[c - 1];
            const int u = P.mean_arow <= 64 ? 1 : P.mean_arow <= 128 ? 2 : P.mean_arow <= 192 ? 3 : 4;
#define SDL_RW1(WS)                                                                          \
    (u == 1 ? k_num_rowacc_win<WS, 1, false, 1> : u == 2 ? k_num_rowacc_win<WS, 2, false, 1> \
                                                : u == 3 ? k_num_rowacc_win<WS, 3, false, 1> \
                                                         : k_num_rowacc_win<WS, 4, false, 1>)
            auto fn = c == 1 ? SDL_RW1(2048) : c == 2 ? SDL_RW1(4096) : SDL_RW1(8192);
#undef SDL_RW1
            const size_t dyn = c == 3 ? 8192 * sizeof(double) : 0;
            hipLaunchKernelGGL(fn, dim3((unsigned) nr), dim3(64), dyn, s, wr, nr, P.L.Lm, P.L.Lp,
                               P.L.Li, P.L.Lperm, nL, P.A.Ap, P.A.Ai, P.Cp.p, P.Ci.p, Lval, d, Ax,
                               Cx);
        }
        long_rows += nwide;
        n1 -= nwide;
        if (P.rank && P.rbin_off[2] - P.rbin_off[1] > nwide)
        {
            // rows of the rank bin; its spills and the longer rows follow
            const int nmid = P.rbin_off[2] - P.rbin_off[1] - nwide;
            const int nwords = (int) ((P.n + 31) / 32);
            const size_t lds = rank_lds_bytes(nwords);
            static bool attr_set = false;
            if (!attr_set)
            {
                SDL_HIP(hipFuncSetAttribute((const void *) k_num_rowacc_rank<1024>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int) rank_lds_bytes(RANK_MAX_WORDS)));
                attr_set = true;
            }
            if (P.rankpos)
            {
                // recorded positions; the spill list is the plan's
                hipLaunchKernelGGL((k_num_rank_pos<RP_G, RP_CH, RP_U>), dim3(P.rank_grid), dim3(RP_G),
                                   (size_t) P.rank_window * sizeof(double), s, P.rrows.p, nmid,
                                   P.rpos.p, P.ipos.p, P.L.Li, P.L.Lperm, P.A.Ap, P.qrange.p,
                                   Lval, d, Ax, Cx, P.rank_window);
            }
            else
            {
                hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, P.nspill.p, 1);
                hipLaunchKernelGGL((k_num_rowacc_rank<1024>), dim3(P.rank_grid), dim3(1024),
                                   lds, s, long_rows, nmid, P.L.Lm, P.L.Lp, P.L.Li, P.L.Lperm,
                                   nL, P.A.Ap, P.A.Ai, P.Cp.p, P.Ci.p, Lval, d, Ax, Cx, nwords,
                                   P.spill.p, P.nspill.p);
            }
            hipLaunchKernelGGL((k_num_rowacc_dense<256>), dim3(P.dense_grid), dim3(256), 0, s,
                               P.spill.p, 0, P.nspill.p, P.L.Lm, P.L.Lp, P.L.Li, P.L.Lperm, nL,
                               P.A.Ap, P.A.Ai, P.Cp.p, P.Ci.p, Lval, d, Ax, Cx, P.acc.p,
                               (int64_t) P.n);
            long_rows += nmid;
            n1 -= nmid;
        }
        if (n1 && P.dense_grid > 0)
            hipLaunchKernelGGL((k_num_rowacc_dense<256>), dim3(P.dense_grid), dim3(256), 0, s,
                               long_rows, n1, (const int *) nullptr, P.L.Lm, P.L.Lp, P.L.Li,
                               P.L.Lperm, nL, P.A.Ap, P.A.Ai, P.Cp.p, P.Ci.p, Lval, d, Ax, Cx,
                               P.acc.p, (int64_t) P.n);
        else if (n1)
            hipLaunchKernelGGL((k_num_rowacc<256, 0>), dim3(n1), dim3(256), 0, s, long_rows,
                               P.L.Lm, P.L.Lp, P.L.Li, P.L.Lperm, nL, P.A.Ap, P.A.Ai, P.Cp.p,
                               P.Ci.p, Lval, d, Ax, Cx);
    }
    SDL_HIP(hipGetLastError());
}

bool product_numeric_multi(const ProductPlan *const *plans, int n, const double *const *Lval,
                           const double *const *d, const double *const *Ax, double *const *Cx,
                           hipStream_t s)
{
    for (int j = 0; j < n; j++)
        if (!plans[j]->pull || !plans[j]->tile_refill || !plans[j]->tq.p) return false;
    for (int j0 = 0; j0 < n; j0 += TN_MAX)
    {
        TileNodes N{};
        int tiles = 0;
        N.n = 0;
        for (int j = j0; j < n && j < j0 + TN_MAX; j++)
        {
            const ProductPlan &P = *plans[j];
            if (P.nnz == 0) continue;
            TileNode &D = N.nd[N.n++];
            D.tqr = P.tqr.p;
            D.ep = P.ep.p;
            D.tq = P.tq.p;
            D.tA = P.tA.p;
            D.Li = P.L.Li;
            D.Lperm = P.L.Lperm;
            D.Lval = Lval[j];
            D.d = d ? d[j] : nullptr;
            D.Ax = Ax[j];
            D.Cx = Cx[j];
            D.nnz = P.nnz;
            D.tile0 = tiles;
            tiles += (int) ((P.nnz + 255) / 256);
        }
        if (N.n == 0) continue;
        hipLaunchKernelGGL(k_num_pull_tile_multi, dim3((unsigned) tiles), dim3(256), 0, s, N);
        SDL_HIP(hipGetLastError());
    }
    return true;
}

void product_bind_values(const ProductPlan &P, const double *Lval, const double *Ax,
                         hipStream_t s)
{
    if (!P.pull || P.terms == 0) return;
    if (P.tv.p)
        hipLaunchKernelGGL(k_gather_terms, dim3((unsigned) ((P.terms + 255) / 256)), dim3(256), 0,
                           s, P.terms, P.tA.p, P.tB.p, Lval, Ax, const_cast<double *>(P.tv.p));
    if (P.ta.p)
    {
        const int64_t n = std::max<int64_t>(P.terms, P.nl);
        hipLaunchKernelGGL(k_bind_tile, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s,
                           P.terms, P.nl, P.tA.p, P.L.Lperm, Lval, Ax,
                           const_cast<double *>(P.ta.p), const_cast<double *>(P.lv.p));
    }
    SDL_HIP(hipGetLastError());
}

void product_numeric_batched(const ProductPlan &P, const double *Lval, const double *D,
                             int64_t ldd, int K, const double *Ax, double *Cx, int64_t ldc,
                             hipStream_t s, const RefillCfg *cfg, unsigned opts, void *ev_start,
                             void *ev_stop, const BatchScratch *ws)
{
    if (P.nnz == 0 || K <= 0) return;
    // workspace: the plan's reservation, or the caller's
    double *tvp = ws ? ws->tv : P.tv.p;
    double *dTp = ws ? ws->dT : P.dT.p;
    const int64_t wld = ws ? ws->ld : P.batch_ld;
    const int cap = ws ? (int) std::min<int64_t>(ws->ld, INT_MAX) : P.batch_cap;
    if (!P.pull)
    {
        for (int r = 0; r < K; r++)
            product_numeric(P, Lval, D + (int64_t) r * ldd, Ax, Cx + (int64_t) r * ldc, s);
        return;
    }
    RefillCfg c = cfg ? *cfg : (P.tuned.variant >= 0 ? P.tuned : default_refill_cfg(P, K));
    const bool kmajor = (opts & BATCH_D_KMAJOR) != 0;
    const bool bound = (opts & BATCH_VALUES_BOUND) != 0;
    // bound values in the plan's tile layout (ta, lv) rather than as pairs
    const bool tileb = bound && !ws && P.ta.p;
    SDL_REQUIRE(!bound || tileb || tvp, "bound refill: no values bound");
    // only the interleaved kernels read a k-major D
    if (kmajor && c.variant < 3) c = RefillCfg{3, c.rchunk == 16 ? 16 : 8};
    SDL_REQUIRE(!kmajor || K <= cap, "k-major refill beyond the reservation");
    if (c.variant >= 3 && K > cap) c = RefillCfg{2, 8};
    // without an operand-pair array (a tile reservation) the interleaved
    // kernels run in tile form
    if (c.variant >= 3 && (tileb || !tvp)) c.variant = 7;
    SDL_REQUIRE(c.variant != 7 || tvp || (P.tq.p && P.tqr.p),
                "batched refill: neither operand pairs nor a tile layout");
    if (c.variant == 7 && K == 1 && !kmajor && P.tq.p)
    {
        // a single refill in tile form: d read in place
        if (ev_start) SDL_HIP(hipEventRecord((hipEvent_t) ev_start, s));
        if (tileb)
            hipLaunchKernelGGL(k_num_pull_tile<true>, dim3((unsigned) ((P.nnz + 255) / 256)),
                               dim3(256), 0, s, P.nnz, P.tqr.p, P.ep.p, P.tq.p, P.tA.p, P.L.Li,
                               P.L.Lperm, P.lv.p, D, P.ta.p, Cx);
        else
            hipLaunchKernelGGL(k_num_pull_tile<false>, dim3((unsigned) ((P.nnz + 255) / 256)),
                               dim3(256), 0, s, P.nnz, P.tqr.p, P.ep.p, P.tq.p, P.tA.p, P.L.Li,
                               P.L.Lperm, Lval, D, Ax, Cx);
        SDL_HIP(hipGetLastError());
        if (ev_stop) SDL_HIP(hipEventRecord((hipEvent_t) ev_stop, s));
        return;
    }
    if (c.variant >= 3 && K == 1 && !kmajor)
    {
        // a single refill: d read in place, no transposition
        if (!(opts & BATCH_VALUES_BOUND) && P.terms > 0)
        {
            hipLaunchKernelGGL(k_gather_terms, dim3((unsigned) ((P.terms + 255) / 256)), dim3(256),
                               0, s, P.terms, P.tA.p, P.tB.p, Lval, Ax, tvp);
            SDL_HIP(hipGetLastError());
        }
        if (ev_start) SDL_HIP(hipEventRecord((hipEvent_t) ev_start, s));
        hipLaunchKernelGGL(k_num_pull_tv1, dim3((unsigned) ((P.nnz + 255) / 256)), dim3(256), 0,
                           s, P.nnz, P.ep.p, P.tK.p, tvp, D, Cx);
        SDL_HIP(hipGetLastError());
        if (ev_stop) SDL_HIP(hipEventRecord((hipEvent_t) ev_stop, s));
        return;
    }
    // direct operand gathers need the caller's value arrays: bound values use
    // the pair array
    if (c.variant == 6 && (opts & BATCH_VALUES_BOUND)) c.variant = 3;
    // the tile form needs the plan's L indices, and the caller's value arrays
    // or values bound in tile layout
    if (c.variant == 7 && (!P.tq.p || !P.tqr.p || (bound && !tileb))) c.variant = 3;
    if (c.variant >= 3)
    {
        const int rc = c.rchunk == 8 ? 8 : 16;
        const int ldt = (int) wld;
        // operand pairs (unless bound) and d transposed to k-major (unless
        // given so), in one launch
        const bool gather = c.variant != 6 && c.variant != 7 && !(opts & BATCH_VALUES_BOUND) &&
                            P.terms > 0;
        const bool transpose = !(opts & BATCH_D_KMAJOR);
        const int ngather = gather ? (int) ((P.terms + 255) / 256) : 0;
        const int ntrans = transpose ? (int) (((P.A.Am + TR_K - 1) / TR_K) *
                                              ((K + TR_R - 1) / TR_R))
                                     : 0;
        if (ngather + ntrans > 0)
        {
            hipLaunchKernelGGL(k_prep_batch, dim3((unsigned) (ngather + ntrans)), dim3(256), 0, s,
                               ngather, P.terms, P.tA.p, P.tB.p, Lval, Ax, tvp, (int) P.A.Am, K,
                               D, ldd, dTp, ldt);
            SDL_HIP(hipGetLastError());
        }
        const double *DT = transpose ? dTp : D;
        const int ldk = transpose ? ldt : (int) ldd;
        if (ev_start) SDL_HIP(hipEventRecord((hipEvent_t) ev_start, s));
        if (c.variant == 7)
        {
            const int ntiles = (int) ((P.nnz + 255) / 256);
            const int nchunks = (K + rc - 1) / rc;
            const int64_t blocks = ((int64_t) ntiles + 7) / 8 * 8 * nchunks;
            const double *lv = tileb ? P.lv.p : Lval;
            const double *av = tileb ? P.ta.p : Ax;
#define SDL_TB(RC_, BV_)                                                                       \
    hipLaunchKernelGGL((k_num_pull_tile_batched<RC_, BV_>), dim3((unsigned) blocks), dim3(256), 0, \
                       s, P.nnz, ntiles, nchunks, P.tqr.p, P.ep.p, P.tq.p, P.tA.p, P.L.Li,      \
                       P.L.Lperm, lv, av, DT, ldk, K, Cx, ldc)
            if (rc == 8 && tileb) SDL_TB(8, true);
            else if (rc == 8) SDL_TB(8, false);
            else if (tileb) SDL_TB(16, true);
            else SDL_TB(16, false);
#undef SDL_TB
            SDL_HIP(hipGetLastError());
            if (ev_stop) SDL_HIP(hipEventRecord((hipEvent_t) ev_stop, s));
            return;
        }
        if (c.variant == 6)
        {
            const int ntiles = (int) ((P.nnz + 255) / 256);
            const int nchunks = (K + rc - 1) / rc;
            const int64_t blocks = ((int64_t) ntiles + 7) / 8 * 8 * nchunks;
            if (rc == 8)
                hipLaunchKernelGGL(k_num_pull_interleaved_g<8>, dim3((unsigned) blocks), dim3(256),
                                   0, s, P.nnz, ntiles, nchunks, P.ep.p, P.tA.p, P.tB.p, P.tK.p,
                                   Lval, Ax, DT, ldk, K, Cx, ldc);
            else
                hipLaunchKernelGGL(k_num_pull_interleaved_g<16>, dim3((unsigned) blocks),
                                   dim3(256), 0, s, P.nnz, ntiles, nchunks, P.ep.p, P.tA.p,
                                   P.tB.p, P.tK.p, Lval, Ax, DT, ldk, K, Cx, ldc);
            SDL_HIP(hipGetLastError());
            if (ev_stop) SDL_HIP(hipEventRecord((hipEvent_t) ev_stop, s));
            return;
        }
        const int ept = c.variant == 4 ? 2 : (c.variant == 5 ? 4 : 1);
        const int ntiles = (int) ((P.nnz + 256 * ept - 1) / (256 * ept));
        const int nchunks = (K + rc - 1) / rc;
        const int64_t blocks = ((int64_t) ntiles + 7) / 8 * 8 * nchunks;
        const bool vec = ((uintptr_t) Cx % (8 * ept) == 0) && (ldc % ept == 0);
#define SDL_IV(RC_, EPT_, VEC_)                                                                 \
    if (rc == RC_ && ept == EPT_ && vec == VEC_)                                              \
    {                                                                                         \
        hipLaunchKernelGGL((k_num_pull_interleaved_v<RC_, EPT_, VEC_>), dim3((unsigned) blocks), \
                           dim3(256), 0, s, P.nnz, ntiles, nchunks, P.ep.p, P.tK.p, tvp,      \
                           DT, ldk, K, Cx, ldc);                                          \
        SDL_HIP(hipGetLastError());                                                           \
        if (ev_stop) SDL_HIP(hipEventRecord((hipEvent_t) ev_stop, s));                       \
        return;                                                                               \
    }
        if (ept > 1)
        {
            SDL_IV(8, 2, true)
            SDL_IV(8, 2, false)
            SDL_IV(16, 2, true)
            SDL_IV(16, 2, false)
            SDL_IV(8, 4, true)
            SDL_IV(8, 4, false)
            SDL_IV(16, 4, true)
            SDL_IV(16, 4, false)
        }
#undef SDL_IV
        if (rc == 8)
            hipLaunchKernelGGL(k_num_pull_interleaved<8>, dim3((unsigned) blocks), dim3(256), 0, s,
                               P.nnz, ntiles, nchunks, P.ep.p, P.tK.p, tvp, DT, ldk, K, Cx, ldc);
        else
            hipLaunchKernelGGL(k_num_pull_interleaved<16>, dim3((unsigned) blocks), dim3(256), 0,
                               s, P.nnz, ntiles, nchunks, P.ep.p, P.tK.p, tvp, DT, ldk, K, Cx,
                               ldc);
        SDL_HIP(hipGetLastError());
        if (ev_stop) SDL_HIP(hipEventRecord((hipEvent_t) ev_stop, s));
        return;
    }
    if (c.variant == 0)
    {
        hipLaunchKernelGGL(k_num_pull_batched, dim3((unsigned) ((P.nnz + 255) / 256)),
                           dim3(256), 0, s, P.nnz, P.ep.p, P.tA.p, P.tB.p, P.tK.p, Lval, D,
                           ldd, K, Ax, Cx, ldc);
        SDL_HIP(hipGetLastError());
        return;
    }
    const int ept = c.variant == 2 ? 2 : 1;
    const int rc = c.rchunk;
    const int64_t per_tile = (int64_t) 256 * ept;
    const int ntiles = (int) ((P.nnz + per_tile - 1) / per_tile);
    const int nchunks = (K + rc - 1) / rc;
    const int64_t tiles_pad = ((int64_t) ntiles + 7) / 8 * 8;
    const int64_t blocks = tiles_pad * nchunks;
    SDL_REQUIRE(blocks < (1ll << 31), "refill grid too large");
#define SDL_RC_CASE(EPT, RC)                                                                 \
    if (ept == EPT && rc == RC)                                                              \
    {                                                                                        \
        hipLaunchKernelGGL((k_num_pull_rchunk<EPT, RC>), dim3((unsigned) blocks), dim3(256), 0, s, \
                           P.nnz, ntiles, nchunks, P.ep.p, P.tA.p, P.tB.p, P.tK.p, Lval, D, ldd, K, \
                           Ax, Cx, ldc);                                                     \
        SDL_HIP(hipGetLastError());                                                          \
        return;                                                                              \
    }
    SDL_RC_CASE(1, 4)
    SDL_RC_CASE(1, 8)
    SDL_RC_CASE(1, 16)
    SDL_RC_CASE(2, 4)
    SDL_RC_CASE(2, 8)
    SDL_RC_CASE(2, 16)
#undef SDL_RC_CASE
    SDL_REQUIRE(false, "unknown refill configuration");
}

RefillCfg default_refill_cfg(const ProductPlan &P, int K)
{
    RefillCfg c;
    if (K <= P.batch_cap)
    {
        // the tile form where the plan has its L indices (bound calls fall
        // back to the operand pairs)
        c.variant = P.tq.p ? 7 : 3;
        c.rchunk = P.tq.p ? 8 : 16;
        return c;
    }
    c.variant = 2;
    c.rchunk = K >= 16 ? 8 : 4;
    return c;
}

int64_t batch_scratch_doubles(const ProductPlan &P, int K)
{
    const int64_t ld = ((int64_t) K + 15) / 16 * 16;
    return 2 * std::max<int64_t>(P.terms, 1) + (K > 1 ? std::max<int64_t>(P.A.Am, 1) * ld : 0);
}

void product_reserve_batch(ProductPlan &P, int Kmax, bool tile)
{
    if (!P.pull) return;
    if (tile)
    {
        SDL_REQUIRE(P.tq.p && P.tqr.p, "tile reservation: the plan has no tile layout");
        if (!P.ta.p)
        {
            int nl = 0;
            d2h(&nl, P.L.Lp + P.L.Lm, sizeof(int));
            P.nl = nl;
            P.ta.alloc((size_t) std::max<int64_t>(P.terms, 1));
            P.lv.alloc((size_t) std::max<int64_t>(P.nl, 1));
        }
    }
    if (Kmax <= P.batch_cap) return;
    if (!P.tv.p && !P.ta.p) P.tv.alloc((size_t) 2 * std::max<int64_t>(P.terms, 1));
    if (Kmax == 1)  // single refills read d in place: no transposed copy
    {
        P.batch_cap = 1;
        return;
    }
    int64_t ld = ((int64_t) Kmax + 15) / 16 * 16;
    P.dT.alloc((size_t) std::max<int64_t>(P.A.Am, 1) * ld);
    P.batch_ld = ld;
    P.batch_cap = (int) ld;
}

void product_numeric_range(const ProductPlan &P, const double *Lval, const double *d,
                           const double *Ax, double *Cx, int64_t e0, int64_t e1, hipStream_t s)
{
    SDL_REQUIRE(P.pull, "range refill: the plan has no pull schedule");
    SDL_REQUIRE(0 <= e0 && e0 <= e1 && e1 <= P.nnz, "range refill: entries out of range");
    if (e1 == e0) return;
    hipLaunchKernelGGL(k_num_pull, dim3((unsigned) ((e1 - e0 + 255) / 256)), dim3(256), 0, s,
                       e1 - e0, P.ep.p + e0, P.tA.p, P.tB.p, P.tK.p, Lval, d, Ax, Cx + e0);
    SDL_HIP(hipGetLastError());
}

std::vector<RefillCfg> refill_candidates()
{
    std::vector<RefillCfg> v;
    RefillCfg c;
    c.variant = 0;
    c.rchunk = 0;
    v.push_back(c);
    for (int var = 1; var <= 2; var++)
        for (int rc : {4, 8, 16})
        {
            c.variant = var;
            c.rchunk = rc;
            v.push_back(c);
        }
    for (int var = 3; var <= 7; var++)
        for (int rc : {8, 16})
        {
            c.variant = var;
            c.rchunk = rc;
            v.push_back(c);
        }
    return v;
}

} // namespace sdl
