// Drop-in matrix sums (reference: include/sparsediff_linalg/matrix_sum.h,
// src/sparse/matrix_sum.c, src/sparse/CSR_sum.c).  The alloc concatenates the
// operands' entries (operand 0 first, each in CSR order) as items, coalesces
// them on the GPU into the union structure and keeps the plan; the fill sums
// each output entry's items in that order from +0.0, so C[r, c] =
// ((+0.0 + x_0) + x_1) + ... over the operands holding (r, c) -- the
// reference's accumulator order (CSR_sum.c:150-157 applied operand by
// operand, matrix_sum.c:142-155; the merge fills CSR_sum.c:84-148 give the
// same sums for canonical operands).  Up to 8 operands with sorted unique
// columns fill through per-entry operand masks (values stream); otherwise
// through the entry -> item lists.
#include "api_matrix.h"
#include "api_pd.h"
#include "sparse_ops.h"

#include <climits>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "../../include/sdl_device.h"

extern "C" {
#include "sparsediff_linalg/matrix_sum.h"
}

namespace sdl {

namespace {

template <typename T> T *bound_as(SparseM *C, const char *fn)
{
    T *b = C->bound ? dynamic_cast<T *>(C->bound.get()) : nullptr;
    if (!b) api_exit(fn, "output was not produced by the matching alloc");
    return b;
}

unsigned sum_grid(int64_t n) { return (unsigned) std::max<int64_t>(1, (n + 255) / 256); }

// Operand owning item t: the largest k with off[k] <= t (empty operands have
// off[k] == off[k+1] and are skipped by taking the largest such k).
__device__ __forceinline__ int op_of(const int64_t *__restrict__ off, int nops, int64_t t)
{
    int lo = 0, hi = nops - 1;
    while (lo < hi)
    {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// map[item] = output entry.
__global__ void k_sum_map(int64_t nnz, const int *__restrict__ seg,
                          const unsigned *__restrict__ order, int *__restrict__ map)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    for (int s = seg[e]; s < seg[e + 1]; s++) map[order[s]] = (int) e;
}

// dup[k] = 1 when two items of operand k land in one entry (repeated (row,
// col) inside operand k).  Items of an entry are in ascending order, so
// repeats of one operand are neighbours.
__global__ void k_sum_dups(int64_t nnz, const int *__restrict__ seg,
                           const unsigned *__restrict__ order, const int64_t *__restrict__ off,
                           int nops, int *__restrict__ dup)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    for (int s = seg[e] + 1; s < seg[e + 1]; s++)
    {
        int k = op_of(off, nops, order[s]);
        if (k == op_of(off, nops, order[s - 1])) dup[k] = 1;
    }
}

// Operand value pointers and item offsets passed by value (up to SUM_ARGS
// operands): no per-call upload, so the fill can be captured into a graph.
constexpr int SUM_ARGS = 32;
struct SumSrc
{
    const double *x[SUM_ARGS];
    int64_t off[SUM_ARGS + 1];
};

// Cx[e] = ((+0.0 + v_1) + v_2) + ... over entry e's items in item order, with
// v = x_k[t - off[k]] or, SCALED, the rounded product d[k*m + row[t]] * x.
// The operand of an item comes from compares against the (wave-uniform)
// offsets, so the only dependent loads are seg -> order -> value.  Items go
// four at a time with unconditional loads (slots past the entry's last item
// repeat its first item and are dropped), so four loads are in flight.
// elist (may be null): the entries to fill, e = elist[i] for i < nnz.
template <bool SCALED>
__global__ __launch_bounds__(256) void k_sum_fill_src(int64_t nnz, const int *__restrict__ seg,
                                                      const unsigned *__restrict__ order,
                                                      const SumSrc src, int nops,
                                                      const int *__restrict__ row,
                                                      const double *__restrict__ d, int m,
                                                      double *__restrict__ Cx,
                                                      const int *__restrict__ elist)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const int64_t e = elist ? elist[i] : i;
    const int s1 = seg[e + 1];
    double acc = 0.0;
    for (int s = seg[e]; s < s1; s += 4)
    {
        const int cnt = s1 - s < 4 ? s1 - s : 4;
        unsigned t[4];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] = order[u < cnt ? s + u : s];
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            const double *base = src.x[0];
            int64_t o = 0;
            int k = 0;
            for (int j = 1; j < nops; j++)
            {
                if ((int64_t) t[u] >= src.off[j])
                {
                    base = src.x[j];
                    o = src.off[j];
                    k = j;
                }
            }
            v[u] = base[t[u] - o];
            if (SCALED) v[u] = d[(int64_t) k * m + row[t[u]]] * v[u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (u < cnt) acc = acc + v[u];
    }
    __builtin_nontemporal_store(acc, Cx + e);
}

// ---- scatter fill (operands that rarely share an output entry) ----
// Output entries fed by one operand entry: Cx[e] = 0.0 + x (or the rounded
// product d * x), the operand entry t streamed in order and written through
// smap[t] = e (-1 for entries of several operand entries, which the list fill
// sums over mlist).
template <bool SCALED>
__global__ __launch_bounds__(256) void k_sum_fill_single(int64_t items,
                                                         const int *__restrict__ smap,
                                                         const SumSrc src, int nops,
                                                         const int *__restrict__ row,
                                                         const double *__restrict__ d, int m,
                                                         double *__restrict__ Cx)
{
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    const int e = smap[t];
    if (e < 0) return;
    const double *base = src.x[0];
    int64_t o = 0;
    int k = 0;
    for (int j = 1; j < nops; j++)
    {
        if (t >= src.off[j])
        {
            base = src.x[j];
            o = src.off[j];
            k = j;
        }
    }
    double v = base[t - o];
    if (SCALED) v = d[(int64_t) k * m + row[t]] * v;
    Cx[e] = 0.0 + v;
}

// smap[t] = map[t] when that entry has one operand entry, else -1.
__global__ void k_sum_single_map(int64_t items, const int *__restrict__ map,
                                 const int *__restrict__ seg, int *__restrict__ smap)
{
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    const int e = map[t];
    smap[t] = seg[e + 1] - seg[e] == 1 ? e : -1;
}

// Operand changes along the output: cnt += entries e whose first operand
// entry belongs to another operand than entry e + 1's.
__global__ void k_sum_switches(int64_t nnz, const int *__restrict__ seg,
                               const unsigned *__restrict__ order,
                               const int64_t *__restrict__ off, int nops,
                               unsigned long long *__restrict__ cnt)
{
    const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const bool sw = e + 1 < nnz &&
                    op_of(off, nops, order[seg[e]]) != op_of(off, nops, order[seg[e + 1]]);
    // one atomic per wavefront
    const unsigned long long b = __ballot(sw);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(cnt, (unsigned long long) __popcll(b));
}

// flag[e] = entry e has several operand entries
__global__ void k_sum_multi_flag(int64_t nnz, const int *__restrict__ seg, int *__restrict__ flag)
{
    const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) flag[e] = seg[e + 1] - seg[e] > 1 ? 1 : 0;
}

__global__ void k_sum_multi_list(int64_t nnz, const int *__restrict__ flag,
                                 const int64_t *__restrict__ pos, int *__restrict__ list)
{
    const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz && flag[e]) list[pos[e]] = (int) e;
}

// ---- operand-mask fill (up to SUM_MASK_OPS operands, sorted unique columns) --
// When every operand's entries land in strictly increasing output entries,
// operand k's value for output entry e is x_k[#(entries before e holding
// operand k)]: the fill needs one mask byte per entry (bit k: operand k holds
// it) and, per SUM_TILE-entry tile (one wave's), each operand's count before
// the tile; the in-tile counts come from wave ballots.  Plan bytes: 1 per entry instead of
// 4 per entry + 4 per operand entry, and the value reads stream.
// Up to 8 operands: one mask byte per entry, 4 entries per lane; 9 to 16:
// two mask bytes per entry, 2 entries per lane (the lane's loads stay at 32).
constexpr int SUM_MASK_OPS = 16;
constexpr int sum_per(int nops) { return nops <= 8 ? 4 : 2; }       // entries per lane
constexpr int sum_tile(int nops) { return 64 * sum_per(nops); }     // entries per wave

template <typename MT>
__global__ void k_sum_masks(int64_t nnz, const int *__restrict__ seg,
                            const unsigned *__restrict__ order, const int64_t *__restrict__ off,
                            int nops, MT *__restrict__ mask)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    unsigned mk = 0;
    for (int s = seg[e]; s < seg[e + 1]; s++) mk |= 1u << op_of(off, nops, order[s]);
    mask[e] = (MT) mk;
}

// bad[k] = 1 when two consecutive entries of operand k do not land in strictly
// increasing output entries (unsorted or repeated columns in a row).
__global__ void k_sum_monotone(int64_t items, const int *__restrict__ map,
                               const int64_t *__restrict__ off, int nops, int *__restrict__ bad)
{
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    const int k = op_of(off, nops, t);
    if (t > off[k] && map[t] <= map[t - 1]) bad[k] = 1;
}

// base[k * ntiles + tile] = operand k's entries in output entries before
// tile * tsize (first j with map_k[j] >= tile * tsize; map_k increases).
__global__ void k_sum_tile_base(int ntiles, int tsize, int nops, const int *__restrict__ map,
                                const int64_t *__restrict__ off, int *__restrict__ base)
{
    int64_t idx = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t) ntiles * nops) return;
    const int k = (int) (idx / ntiles), tile = (int) (idx % ntiles);
    const int64_t e0 = (int64_t) tile * tsize;
    int64_t lo = off[k], hi = off[k + 1];
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) >> 1;
        if (map[mid] < e0) lo = mid + 1;
        else hi = mid;
    }
    base[idx] = (int) (lo - off[k]);
}

// One wave per tile of PER * 64 entries: PER rows of 64, lane i holding
// entry i of each (every load and store coalesced), no LDS and no barrier.
// Cx[e] = ((+0.0 + x_k0) + x_k1) + ... over the operands in e's mask, ascending
// k.  NOPS is the operand count, so the loops have no branches and every value
// load is issued unconditionally (a lane without operand k reads the mask
// buffer's first word instead and drops it): all SUM_PER * NOPS loads of a
// lane are in flight together.
// SCALED: each term is the rounded product d[k*m + row] * x, the row read from
// erow (one int per entry).
template <int NOPS, bool SCALED, typename MT, int PER>
__global__ __launch_bounds__(256) void k_sum_fill_mask(int64_t nnz, int ntiles,
                                                      const MT *__restrict__ mask,
                                                      const int *__restrict__ base,
                                                      const SumSrc src,
                                                      const int *__restrict__ erow,
                                                      const double *__restrict__ d, int m,
                                                      double *__restrict__ Cx)
{
    const int lane = threadIdx.x & 63;
    const int tile = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (tile >= ntiles) return;  // whole waves
    const int64_t t0 = (int64_t) tile * (64 * PER) + lane;
    const unsigned long long lt = (1ull << lane) - 1ull;
    const double *dummy = (const double *) mask;
    unsigned mk[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) mk[u] = mask[t0 + 64 * u];  // padded to whole tiles
    int run[NOPS];
#pragma unroll
    for (int k = 0; k < NOPS; k++) run[k] = base[(int64_t) k * ntiles + tile];
    double v[PER][NOPS];
#pragma unroll
    for (int u = 0; u < PER; u++)
    {
#pragma unroll
        for (int k = 0; k < NOPS; k++)
        {
            const bool hit = (mk[u] >> k) & 1u;
            const unsigned long long b = __ballot(hit);
            const double *a = hit ? src.x[k] + run[k] + __popcll(b & lt) : dummy;
            v[u][k] = *a;
            run[k] += __popcll(b);
        }
    }
    if (SCALED)
    {
#pragma unroll
        for (int u = 0; u < PER; u++)
        {
            const int64_t e = t0 + 64 * u;
            const int64_t r = erow[e < nnz ? e : nnz - 1];
#pragma unroll
            for (int k = 0; k < NOPS; k++) v[u][k] = d[(int64_t) k * m + r] * v[u][k];
        }
    }
#pragma unroll
    for (int u = 0; u < PER; u++)
    {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < NOPS; k++)
            if ((mk[u] >> k) & 1u) acc = acc + v[u][k];
        const int64_t e = t0 + 64 * u;
        if (e < nnz) __builtin_nontemporal_store(acc, Cx + e);
    }
}

template <int NOPS>
void launch_fill_mask(unsigned blocks, hipStream_t st, int64_t nnz, int ntiles,
                      const void *mask, const int *base, const SumSrc &src,
                      const int *erow, const double *d, int m, double *Cx)
{
    typedef typename std::conditional<(NOPS <= 8), unsigned char, unsigned short>::type MT;
    constexpr int PER = sum_per(NOPS);
    const MT *mk = (const MT *) mask;
    if (d)
        hipLaunchKernelGGL((k_sum_fill_mask<NOPS, true, MT, PER>), dim3(blocks), dim3(256), 0,
                           st, nnz, ntiles, mk, base, src, erow, d, m, Cx);
    else
        hipLaunchKernelGGL((k_sum_fill_mask<NOPS, false, MT, PER>), dim3(blocks), dim3(256), 0,
                           st, nnz, ntiles, mk, base, src, erow, d, m, Cx);
}

// erow[e] = row of output entry e (the largest r with Cp[r] <= e).
__global__ void k_sum_entry_rows(int64_t nnz, int m, const int64_t *__restrict__ Cp,
                                 int *__restrict__ erow)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    int lo = 0, hi = m - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (Cp[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    erow[e] = lo;
}

// The same fill with the pointers and offsets in device memory (any count).
template <bool SCALED>
__global__ void k_sum_fill(int64_t nnz, const int *__restrict__ seg,
                           const unsigned *__restrict__ order, const int64_t *__restrict__ off,
                           int nops, const double *const *__restrict__ xs,
                           const int *__restrict__ row, const double *__restrict__ d, int m,
                           double *__restrict__ Cx)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double acc = 0.0;
    for (int s = seg[e]; s < seg[e + 1]; s++)
    {
        const unsigned t = order[s];
        const int k = op_of(off, nops, t);
        double v = xs[k][t - off[k]];
        if (SCALED) v = d[(int64_t) k * m + row[t]] * v;
        acc = acc + v;
    }
    Cx[e] = acc;
}

// Cx[map[j]] += x[j] for an operand without repeated entries (its items land
// in distinct entries, so the adds do not collide).
__global__ void k_sum_add(int64_t cnt, const int *__restrict__ map, const double *__restrict__ x,
                          double *__restrict__ Cx)
{
    int64_t j = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (j < cnt) Cx[map[j]] = Cx[map[j]] + x[j];
}

// The same add for an operand with repeated entries: each entry adds its items
// of operand [lo, hi) in item order.
__global__ void k_sum_add_ordered(int64_t nnz, const int *__restrict__ seg,
                                  const unsigned *__restrict__ order, int64_t lo, int64_t hi,
                                  const double *__restrict__ x, double *__restrict__ Cx)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double acc = Cx[e];
    bool hit = false;
    for (int s = seg[e]; s < seg[e + 1]; s++)
    {
        const int64_t t = order[s];
        if (t >= lo && t < hi)
        {
            acc = acc + x[t - lo];
            hit = true;
        }
    }
    if (hit) Cx[e] = acc;
}

// Union plan of nops same-shape operands (device CSR structures).
struct SumPlan
{
    int nops = 0, m = 0, n = 0;
    CoalescePlan P;                 // structure; its int64 segments are narrowed into seg
    DBuf<int> seg;                  // nnz + 1: items of entry e are order[seg[e] .. seg[e+1])
    std::vector<int64_t> off;       // nops + 1 item offsets
    DBuf<int64_t> doff;
    DBuf<int> map;                  // item -> output entry
    DBuf<int> row;                  // item -> row (kept for the scaled fill)
    std::vector<int> dup;           // operand k repeats an entry
    bool masked = false;            // the operand-mask fill applies
    int ntiles = 0;
    DBuf<unsigned char> mask;       // operand masks, 1 byte (<= 8 operands) or 2 per entry
    DBuf<int> tile_base;            // nops x ntiles (masked plans)
    DBuf<int> erow;                 // row of every output entry (masked plans keeping rows)
    DBuf<const double *> xs;        // value pointers beyond SUM_ARGS operands
    bool scatter = false;           // the scatter fill applies (few shared entries)
    DBuf<int> smap;                 // item -> its entry if single-source, else -1
    DBuf<int> mlist;                // entries of several items (scatter plans)
    int64_t nmulti = 0;
};

double g_sum_scatter_ratio = 1.25;
// mean run of consecutive output entries from one operand the scatter fill
// needs (16 doubles: 128 bytes, one line)
constexpr double SUM_SCATTER_RUN = 16.0;
// ... or an output small enough (32 MB: the eight XCDs' L2) that the caches
// merge its scattered writes
constexpr int64_t SUM_SCATTER_CACHED = 4 << 20;

void sum_plan_build(SumPlan &S, int nops, int m, int n, const int *const *dP,
                    const int *const *dI, const int64_t *nnz, bool keep_rows)
{
    S.nops = nops;
    S.m = m;
    S.n = n;
    S.off.assign((size_t) nops + 1, 0);
    for (int k = 0; k < nops; k++) S.off[k + 1] = S.off[k] + nnz[k];
    const int64_t items = S.off[nops];
    DBuf<int> orow(items), ocol(items);
    for (int k = 0; k < nops; k++)
    {
        if (nnz[k] == 0) continue;
        expand_rows(m, dP[k], orow.p + S.off[k], stream());
        d2d(ocol.p + S.off[k], dI[k], (size_t) nnz[k] * sizeof(int));
    }
    coalesce_plan(S.P, items, orow.p, ocol.p, m, n);
    S.seg.alloc((size_t) S.P.nnz + 1);
    narrow_i64(S.P.nnz + 1, S.P.seg.p, S.seg.p, stream());
    S.doff = upload(S.off.data(), nops + 1);
    S.map.alloc(items);
    DBuf<int> ddup(nops);
    dzero(ddup.p, ddup.bytes());
    if (S.P.nnz > 0)
    {
        hipLaunchKernelGGL(k_sum_map, dim3(sum_grid(S.P.nnz)), dim3(256), 0, stream(), S.P.nnz,
                           S.seg.p, S.P.order.p, S.map.p);
        hipLaunchKernelGGL(k_sum_dups, dim3(sum_grid(S.P.nnz)), dim3(256), 0, stream(), S.P.nnz,
                           S.seg.p, S.P.order.p, S.doff.p, nops, ddup.p);
    }
    SDL_HIP(hipGetLastError());
    S.dup.assign(nops, 0);
    d2h(S.dup.data(), ddup.p, (size_t) nops * sizeof(int));
    S.P.seg.reset();
    // operands that rarely share an output entry and each cover long runs of
    // consecutive output entries (so their scattered writes still fill whole
    // lines): the scatter fill
    bool scatter = nops <= SUM_ARGS && S.P.nnz > 0 &&
                   (double) items <= g_sum_scatter_ratio * (double) S.P.nnz;
    if (scatter && nops > 1)
    {
        DBuf<unsigned long long> sw(1);
        dzero(sw.p, sw.bytes());
        hipLaunchKernelGGL(k_sum_switches, dim3(sum_grid(S.P.nnz)), dim3(256), 0, stream(),
                           S.P.nnz, S.seg.p, S.P.order.p, S.doff.p, nops, sw.p);
        SDL_HIP(hipGetLastError());
        unsigned long long h = 0;
        d2h(&h, sw.p, sizeof(h));
        scatter = (double) S.P.nnz >= SUM_SCATTER_RUN * (double) (h + 1) ||
                  S.P.nnz <= SUM_SCATTER_CACHED;
    }
    if (scatter)
    {
        S.scatter = true;
        S.smap.alloc(items);
        hipLaunchKernelGGL(k_sum_single_map, dim3(sum_grid(items)), dim3(256), 0, stream(), items,
                           S.map.p, S.seg.p, S.smap.p);
        DBuf<int> flag(S.P.nnz);
        DBuf<int64_t> pos((size_t) S.P.nnz + 1);
        hipLaunchKernelGGL(k_sum_multi_flag, dim3(sum_grid(S.P.nnz)), dim3(256), 0, stream(),
                           S.P.nnz, S.seg.p, flag.p);
        S.nmulti = scan_i32_to_i64(flag.p, pos.p, S.P.nnz, true);
        S.mlist.alloc((size_t) std::max<int64_t>(S.nmulti, 1));
        if (S.nmulti > 0)
            hipLaunchKernelGGL(k_sum_multi_list, dim3(sum_grid(S.P.nnz)), dim3(256), 0, stream(),
                               S.P.nnz, flag.p, pos.p, S.mlist.p);
        SDL_HIP(hipGetLastError());
        sync();
    }
    if (!S.scatter && nops <= SUM_MASK_OPS && S.P.nnz > 0)
    {
        DBuf<int> dbad(nops);
        dzero(dbad.p, dbad.bytes());
        hipLaunchKernelGGL(k_sum_monotone, dim3(sum_grid(items)), dim3(256), 0, stream(), items,
                           S.map.p, S.doff.p, nops, dbad.p);
        SDL_HIP(hipGetLastError());
        std::vector<int> bad(nops, 0);
        d2h(bad.data(), dbad.p, (size_t) nops * sizeof(int));
        S.masked = true;
        for (int k = 0; k < nops; k++) S.masked = S.masked && !bad[k];
    }
    if (S.masked)
    {
        const int tsize = sum_tile(nops);
        const size_t mbytes = nops <= 8 ? 1 : 2;
        S.ntiles = (int) ((S.P.nnz + tsize - 1) / tsize);
        // padded to whole tiles (zero masks)
        S.mask.alloc((size_t) S.ntiles * tsize * mbytes);
        dzero(S.mask.p, S.mask.bytes());
        S.tile_base.alloc((size_t) S.ntiles * nops);
        if (mbytes == 1)
            hipLaunchKernelGGL(k_sum_masks<unsigned char>, dim3(sum_grid(S.P.nnz)), dim3(256), 0,
                               stream(), S.P.nnz, S.seg.p, S.P.order.p, S.doff.p, nops,
                               S.mask.p);
        else
            hipLaunchKernelGGL(k_sum_masks<unsigned short>, dim3(sum_grid(S.P.nnz)), dim3(256),
                               0, stream(), S.P.nnz, S.seg.p, S.P.order.p, S.doff.p, nops,
                               (unsigned short *) S.mask.p);
        hipLaunchKernelGGL(k_sum_tile_base, dim3(sum_grid((int64_t) S.ntiles * nops)), dim3(256),
                           0, stream(), S.ntiles, tsize, nops, S.map.p, S.doff.p,
                           S.tile_base.p);
        SDL_HIP(hipGetLastError());
    }
    if (S.masked && keep_rows)
    {
        S.erow.alloc((size_t) S.P.nnz);
        hipLaunchKernelGGL(k_sum_entry_rows, dim3(sum_grid(S.P.nnz)), dim3(256), 0, stream(),
                           S.P.nnz, m, S.P.Cp.p, S.erow.p);
        SDL_HIP(hipGetLastError());
    }
    if (keep_rows) S.row = std::move(orow);
    if (nops > SUM_ARGS) S.xs.alloc(nops);
}

// dCx = the sum of the operands whose device values are hX[k] (host array of
// device pointers); dd (nops x m row scales, device) gives the scaled sum.
// Up to SUM_ARGS operands the launch carries the pointers and allocates and
// uploads nothing.
void sum_plan_fill(SumPlan &S, const double *const *hX, const double *dd, double *dCx,
                   hipStream_t st)
{
    if (S.P.nnz <= 0) return;
    const unsigned grid = sum_grid(S.P.nnz);
    if (S.nops <= SUM_ARGS)
    {
        SumSrc src{};
        for (int k = 0; k < S.nops; k++) src.x[k] = hX[k];
        for (int k = 0; k <= S.nops; k++) src.off[k] = S.off[k];
        if (S.scatter)
        {
            const int64_t items = S.off[S.nops];
            if (dd)
                hipLaunchKernelGGL(k_sum_fill_single<true>, dim3(sum_grid(items)), dim3(256), 0, st,
                                   items, S.smap.p, src, S.nops, S.row.p, dd, S.m, dCx);
            else
                hipLaunchKernelGGL(k_sum_fill_single<false>, dim3(sum_grid(items)), dim3(256), 0,
                                   st, items, S.smap.p, src, S.nops, (const int *) nullptr,
                                   (const double *) nullptr, S.m, dCx);
            if (S.nmulti > 0)
            {
                if (dd)
                    hipLaunchKernelGGL(k_sum_fill_src<true>, dim3(sum_grid(S.nmulti)), dim3(256), 0,
                                       st, S.nmulti, S.seg.p, S.P.order.p, src, S.nops, S.row.p,
                                       dd, S.m, dCx, S.mlist.p);
                else
                    hipLaunchKernelGGL(k_sum_fill_src<false>, dim3(sum_grid(S.nmulti)), dim3(256),
                                       0, st, S.nmulti, S.seg.p, S.P.order.p, src, S.nops,
                                       (const int *) nullptr, (const double *) nullptr, S.m, dCx,
                                       S.mlist.p);
            }
        }
        else if (S.masked)
        {
            const unsigned blocks = (unsigned) ((S.ntiles + 3) / 4);
            auto *fn = &launch_fill_mask<1>;
            switch (S.nops)
            {
            case 2: fn = &launch_fill_mask<2>; break;
            case 3: fn = &launch_fill_mask<3>; break;
            case 4: fn = &launch_fill_mask<4>; break;
            case 5: fn = &launch_fill_mask<5>; break;
            case 6: fn = &launch_fill_mask<6>; break;
            case 7: fn = &launch_fill_mask<7>; break;
            case 8: fn = &launch_fill_mask<8>; break;
            case 9: fn = &launch_fill_mask<9>; break;
            case 10: fn = &launch_fill_mask<10>; break;
            case 11: fn = &launch_fill_mask<11>; break;
            case 12: fn = &launch_fill_mask<12>; break;
            case 13: fn = &launch_fill_mask<13>; break;
            case 14: fn = &launch_fill_mask<14>; break;
            case 15: fn = &launch_fill_mask<15>; break;
            case 16: fn = &launch_fill_mask<16>; break;
            default: break;
            }
            fn(blocks, st, S.P.nnz, S.ntiles, S.mask.p, S.tile_base.p, src, S.erow.p, dd, S.m,
               dCx);
        }
        else if (dd)
            hipLaunchKernelGGL(k_sum_fill_src<true>, dim3(grid), dim3(256), 0, st, S.P.nnz,
                               S.seg.p, S.P.order.p, src, S.nops, S.row.p, dd, S.m, dCx,
                               (const int *) nullptr);
        else
            hipLaunchKernelGGL(k_sum_fill_src<false>, dim3(grid), dim3(256), 0, st, S.P.nnz,
                               S.seg.p, S.P.order.p, src, S.nops, (const int *) nullptr,
                               (const double *) nullptr, S.m, dCx, (const int *) nullptr);
    }
    else
    {
        h2d(S.xs.p, hX, (size_t) S.nops * sizeof(const double *));
        if (dd)
            hipLaunchKernelGGL(k_sum_fill<true>, dim3(grid), dim3(256), 0, st, S.P.nnz, S.seg.p,
                               S.P.order.p, S.doff.p, S.nops, S.xs.p, S.row.p, dd, S.m, dCx);
        else
            hipLaunchKernelGGL(k_sum_fill<false>, dim3(grid), dim3(256), 0, st, S.P.nnz,
                               S.seg.p, S.P.order.p, S.doff.p, S.nops, S.xs.p,
                               (const int *) nullptr, (const double *) nullptr, S.m, dCx);
    }
    SDL_HIP(hipGetLastError());
}

// dCx += operand k (device values x), in the reference's accumulator order.
void sum_plan_add(SumPlan &S, int k, const double *x, double *dCx, hipStream_t st)
{
    const int64_t lo = S.off[k], hi = S.off[k + 1];
    if (hi <= lo) return;
    if (!S.dup[k])
        hipLaunchKernelGGL(k_sum_add, dim3(sum_grid(hi - lo)), dim3(256), 0, st, hi - lo,
                           S.map.p + lo, x, dCx);
    else
        hipLaunchKernelGGL(k_sum_add_ordered, dim3(sum_grid(S.P.nnz)), dim3(256), 0, st, S.P.nnz,
                           S.seg.p, S.P.order.p, lo, hi, x, dCx);
    SDL_HIP(hipGetLastError());
}

struct SumBound : Bound
{
    SumPlan S;
    std::vector<const double *> hxs;
    DBuf<double> d;                 // 2 x m row scales (scaled fill)
};

// Plan over the operands' CSR views (any kinds).
std::unique_ptr<SumBound> sum_plan(const char *fn, matrix *const *ops, int nops, bool keep_rows)
{
    if (nops < 1) api_exit(fn, "need at least one operand");
    const int m = ops[0]->m, n = ops[0]->n;
    std::vector<const int *> dP(nops), dI(nops);
    std::vector<int64_t> nnz(nops);
    int64_t items = 0;
    for (int k = 0; k < nops; k++)
    {
        if (ops[k]->m != m || ops[k]->n != n) api_exit(fn, "operand shapes differ");
        SparseM *S = as_sparse(pd_csr_operand(ops[k]));
        sparse_ensure_dev(S);
        dP[k] = S->dp.p;
        dI[k] = S->di.p;
        nnz[k] = ops[k]->nnz;
        items += nnz[k];
    }
    if (items > INT_MAX) api_exit(fn, "total operand nnz exceeds INT_MAX.");
    auto b = std::make_unique<SumBound>();
    sum_plan_build(b->S, nops, m, n, dP.data(), dI.data(), nnz.data(), keep_rows);
    if (keep_rows) b->d.alloc((size_t) 2 * m);
    b->hxs.assign(nops, nullptr);
    return b;
}

// Row pointers and columns of the plan into a host CSR with room for them.
void write_structure(CSR_matrix *c, const CoalescePlan &P)
{
    DBuf<int> p32((size_t) c->m + 1);
    narrow_i64((int64_t) c->m + 1, P.Cp.p, p32.p, stream());
    d2h(c->p, p32.p, ((size_t) c->m + 1) * sizeof(int));
    if (P.nnz > 0) d2h(c->i, P.Ci.p, (size_t) P.nnz * sizeof(int));
    c->nnz = (int) P.nnz;
}

matrix *sum_output(const char *fn, std::unique_ptr<SumBound> b, int m, int n)
{
    const int nnz = checked_nnz(fn, b->S.P.nnz);
    CSR_matrix *c = new_CSR_matrix(m, n, nnz);
    write_structure(c, b->S.P);
    SparseM *C = as_sparse(make_sparse(c));
    sparse_ensure_dev(C);
    C->bound = std::move(b);
    return &C->base;
}

void check_operands(const char *fn, const SumBound *b, const matrix *const *ops, int nops)
{
    if (b->S.nops != nops) api_exit(fn, "operand count differs from the alloc");
    for (int k = 0; k < nops; k++)
        if (ops[k]->nnz != b->S.off[k + 1] - b->S.off[k])
            api_exit(fn, "operand nnz differs from the alloc");
}

void sum_fill(const char *fn, matrix *const *ops, int nops, matrix *Cm, const double *d1,
              const double *d2)
{
    SparseM *C = as_sparse(Cm);
    SumBound *b = bound_as<SumBound>(C, fn);
    check_operands(fn, b, ops, nops);
    for (int k = 0; k < nops; k++)
        b->hxs[k] = ops[k]->nnz > 0 ? sparse_dev_values(as_sparse(pd_csr_operand_fresh(ops[k])))
                                    : nullptr;
    const double *dd = nullptr;
    if (d1)
    {
        if (!b->S.row.p) api_exit(fn, "output was not produced by sum_matrices_alloc");
        h2d(b->d.p, d1, (size_t) b->S.m * sizeof(double));
        h2d(b->d.p + b->S.m, d2, (size_t) b->S.m * sizeof(double));
        dd = b->d.p;
    }
    sum_plan_fill(b->S, b->hxs.data(), dd, C->val.x.p, stream());
    sparse_publish(C);
}

} // namespace

} // namespace sdl

using namespace sdl;

struct sdl_sum_plan
{
    SumPlan S;
};

extern "C" {
#pragma GCC visibility push(default)

void sum_matrices_alloc(matrix *A, matrix *B, matrix *Cm)
{
    const char *fn = "sum_matrices_alloc";
    if (Cm->kind != MATRIX_SPARSE) api_exit(fn, "C must be a sparse matrix");
    SparseM *C = as_sparse(Cm);
    if (C->ext_values) api_exit(fn, "C must own its values");
    if (Cm->m != A->m || Cm->n != A->n) api_exit(fn, "C's shape differs from the operands'");
    matrix *ops[2] = {A, B};
    std::unique_ptr<SumBound> b = sum_plan(fn, ops, 2, true);
    const int nnz = checked_nnz(fn, b->S.P.nnz);
    // C's CSR keeps its struct and row-pointer array; column and value arrays
    // are sized to the union (the reference's CSR_trim, CSR_matrix.c:56-63).
    CSR_matrix *c = C->csr;
    detach(c);
    const size_t cap = (size_t) std::max(nnz, 1);
    int *ni = (int *) std::realloc(c->i, cap * sizeof(int));
    double *nx = (double *) std::realloc(c->x, cap * sizeof(double));
    if (!ni || !nx) api_exit(fn, "host allocation failed");
    c->i = ni;
    c->x = nx;
    write_structure(c, b->S.P);
    Cm->x = c->x;
    Cm->nnz = nnz;
    // the device mirror follows the new structure
    sdl::sync();
    C->dev = false;
    C->dp.reset();
    C->di.reset();
    C->val.x.reset();
    C->val.valid = false;
    C->drow.reset();
    C->has_csc = false;
    C->cp.reset();
    C->ci.reset();
    C->cperm.reset();
    sparse_ensure_dev(C);
    C->bound = std::move(b);
}

void sum_matrices_fill_values(matrix *A, matrix *B, matrix *C)
{
    matrix *ops[2] = {A, B};
    sum_fill("sum_matrices_fill_values", ops, 2, C, nullptr, nullptr);
}

void sum_scaled_matrices_fill_values(matrix *A, matrix *B, matrix *C, const double *d1,
                                     const double *d2)
{
    matrix *ops[2] = {A, B};
    sum_fill("sum_scaled_matrices_fill_values", ops, 2, C, d1, d2);
}

matrix *sum_4_matrices_alloc(matrix *A, matrix *B, matrix *D, matrix *E)
{
    matrix *ops[4] = {A, B, D, E};
    return sum_output("sum_4_matrices_alloc", sum_plan("sum_4_matrices_alloc", ops, 4, false),
                      A->m, A->n);
}

void sum_4_matrices_fill_values(const matrix *A, const matrix *B, const matrix *D,
                                const matrix *E, matrix *C)
{
    matrix *ops[4] = {const_cast<matrix *>(A), const_cast<matrix *>(B), const_cast<matrix *>(D),
                      const_cast<matrix *>(E)};
    sum_fill("sum_4_matrices_fill_values", ops, 4, C, nullptr, nullptr);
}

matrix *sum_n_matrices_alloc(matrix *const *ops, int n)
{
    if (n < 1) api_exit("sum_n_matrices_alloc", "need at least one operand");
    return sum_output("sum_n_matrices_alloc", sum_plan("sum_n_matrices_alloc", ops, n, false),
                      ops[0]->m, ops[0]->n);
}

void sum_n_matrices_fill_values(matrix *const *ops, int n, matrix *C)
{
    sum_fill("sum_n_matrices_fill_values", ops, n, C, nullptr, nullptr);
}

void sum_n_matrices_fill_zero(matrix *Cm)
{
    SparseM *C = as_sparse(Cm);
    (void) bound_as<SumBound>(C, "sum_n_matrices_fill_zero");
    if (Cm->nnz > 0)
    {
        std::memset(C->csr->x, 0, (size_t) Cm->nnz * sizeof(double));
        dzero(C->val.x.p, (size_t) Cm->nnz * sizeof(double));
    }
    matrix_values_changed(Cm);
    C->val.seen = Cm->values_version;
    C->val.valid = true;
}

void sum_n_matrices_fill_add(const matrix *op, int k, matrix *Cm)
{
    const char *fn = "sum_n_matrices_fill_add";
    SparseM *C = as_sparse(Cm);
    SumBound *b = bound_as<SumBound>(C, fn);
    if (k < 0 || k >= b->S.nops) api_exit(fn, "operand index out of range");
    if (op->nnz != b->S.off[k + 1] - b->S.off[k]) api_exit(fn, "operand nnz differs from the alloc");
    if (op->nnz > 0)
    {
        // C's device values current (the caller may have written C->x)
        double *cx = const_cast<double *>(sparse_dev_values(C));
        const double *x =
            sparse_dev_values(as_sparse(pd_csr_operand_fresh(const_cast<matrix *>(op))));
        sum_plan_add(b->S, k, x, cx, stream());
    }
    sparse_publish(C);
}

// ---- device-resident sums ----
int sdl_sum_plan_create(int nops, int m, int n, const int *const *dP, const int *const *dI,
                        const long long *nnz, int keep_rows, sdl_sum_plan **out)
{
    if (!out || nops < 1 || m < 0 || n < 0 || !dP || !dI || !nnz)
    {
        set_error("sdl_sum_plan_create: invalid arguments");
        return 1;
    }
    long long items = 0;
    for (int k = 0; k < nops; k++)
    {
        if (nnz[k] < 0 || (nnz[k] > 0 && (!dP[k] || !dI[k])) || (m > 0 && !dP[k]))
        {
            set_error("sdl_sum_plan_create: operand %d: invalid arguments", k);
            return 1;
        }
        items += nnz[k];
    }
    if (items > INT_MAX)
    {
        set_error("sdl_sum_plan_create: total operand nnz exceeds INT_MAX");
        return 1;
    }
    sdl_sum_plan *p = new (std::nothrow) sdl_sum_plan();
    if (!p) return 1;
    std::vector<int64_t> nz(nnz, nnz + nops);
    sum_plan_build(p->S, nops, m, n, dP, dI, nz.data(), keep_rows != 0);
    *out = p;
    return 0;
}

long long sdl_sum_plan_nnz(const sdl_sum_plan *p) { return p ? p->S.P.nnz : -1; }
const long long *sdl_sum_plan_rowptr(const sdl_sum_plan *p)
{
    return p ? (const long long *) p->S.P.Cp.p : nullptr;
}
const int *sdl_sum_plan_colind(const sdl_sum_plan *p) { return p ? p->S.P.Ci.p : nullptr; }
const int *sdl_sum_plan_map(const sdl_sum_plan *p) { return p ? p->S.map.p : nullptr; }
int sdl_sum_plan_is_masked(const sdl_sum_plan *p) { return p && p->S.masked ? 1 : 0; }

int sdl_sum_plan_fill_kind(const sdl_sum_plan *p)
{
    if (!p) return -1;
    return p->S.scatter ? 2 : (p->S.masked ? 1 : 0);
}

double sdl_sum_scatter_ratio(double ratio)
{
    const double prev = sdl::g_sum_scatter_ratio;
    if (ratio >= 0.0) sdl::g_sum_scatter_ratio = ratio;
    return prev;
}

int sdl_sum_fill(sdl_sum_plan *p, const double *const *dX, const double *dD, double *dCx,
                 void *stream)
{
    if (!p || (p->S.P.nnz > 0 && (!dX || !dCx)))
    {
        set_error("sdl_sum_fill: invalid arguments");
        return 1;
    }
    if (dD && !p->S.row.p)
    {
        set_error("sdl_sum_fill: row scales need a plan created with keep_rows");
        return 1;
    }
    sum_plan_fill(p->S, dX, dD, dCx, stream ? (hipStream_t) stream : sdl::stream());
    return 0;
}

int sdl_sum_add(sdl_sum_plan *p, int k, const double *dx, double *dCx, void *stream)
{
    if (!p || k < 0 || k >= p->S.nops)
    {
        set_error("sdl_sum_add: invalid arguments");
        return 1;
    }
    sum_plan_add(p->S, k, dx, dCx, stream ? (hipStream_t) stream : sdl::stream());
    return 0;
}

void sdl_sum_plan_destroy(sdl_sum_plan *p)
{
    if (!p) return;
    sdl::sync();
    delete p;
}

#pragma GCC visibility pop
} // extern "C"
