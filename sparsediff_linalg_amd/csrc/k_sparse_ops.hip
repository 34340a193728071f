// Sparse-format operations on the device.  Transposition and coalescing are
// both a stable radix sort of entry keys in source order: stability is what
// reproduces the reference's counting-sort order (CSC_matrix.c:62-123) and its
// input-order duplicate sums (sparse_matrix.c:136-240, CSR_sum.c:150-157).
#include "sparse_ops.h"

#include <algorithm>

namespace sdl {

namespace {

inline unsigned grid_for(int64_t n, int block = 256)
{
    return (unsigned) std::max<int64_t>(1, (n + block - 1) / block);
}

inline int bits_for(int64_t range)
{
    int b = 0;
    while (b < 63 && ((int64_t) 1 << b) < range) b++;
    return b;
}

__global__ void k_expand_rows(int m, const int *__restrict__ Ap, int *__restrict__ rowidx)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    for (int t = Ap[r]; t < Ap[r + 1]; t++) rowidx[t] = r;
}

template <typename K>
__global__ void k_bounds_from_sorted(const K *__restrict__ keys, int64_t nnz, int64_t nb,
                                     int shift, int64_t *__restrict__ out64,
                                     int *__restrict__ out32)
{
    int64_t c = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nb) return;
    K target = (K) c << shift;
    int64_t lo = 0, hi = nnz;
    while (lo < hi)
    {
        int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    if (out64) out64[c] = lo;
    if (out32) out32[c] = (int) lo;
}

__global__ void k_copy_keys_u32(int64_t n, const int *__restrict__ in, unsigned *__restrict__ out)
{
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (unsigned) in[t];
}

__global__ void k_transpose_finish(int64_t nnz, const unsigned *__restrict__ sorted_src,
                                   const int *__restrict__ rowidx, int *__restrict__ ATi,
                                   int *__restrict__ perm)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    unsigned src = sorted_src[e];
    ATi[e] = rowidx[src];
    perm[e] = (int) src;
}

__global__ void k_gather_values(int64_t n, const int *__restrict__ perm,
                                const double *__restrict__ in, double *__restrict__ out)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = in[perm[e]];
}

template <typename K>
__global__ void k_make_keys(int64_t n, const int *__restrict__ orow, const int *__restrict__ ocol,
                            int cbits, K *__restrict__ keys)
{
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) keys[t] = ((K) (unsigned) orow[t] << cbits) | (K) (unsigned) ocol[t];
}

template <typename K>
__global__ void k_heads(int64_t n, const K *__restrict__ keys, int *__restrict__ head)
{
    int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) head[s] = (s == 0 || keys[s] != keys[s - 1]) ? 1 : 0;
}

template <typename K>
__global__ void k_coalesce_entries(int64_t n, const K *__restrict__ keys,
                                   const int *__restrict__ head, const int64_t *__restrict__ hoff,
                                   K cmask, int *__restrict__ Ci, int64_t *__restrict__ seg,
                                   K *__restrict__ ekeys)
{
    int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || !head[s]) return;
    int64_t e = hoff[s];
    Ci[e] = (int) (keys[s] & cmask);
    seg[e] = s;
    ekeys[e] = keys[s];
}

// Cx[e] = ((+0.0 + x[t1]) + x[t2]) + ... over entry e's items in order.  Items
// go four at a time with unconditional loads (indices past the entry's last
// item repeat its first, whose value is then dropped), so up to four random
// gathers of x are in flight per thread instead of one.
__global__ __launch_bounds__(256) void k_coalesce_fill(int64_t nnz,
                                                      const int64_t *__restrict__ seg,
                                                      const unsigned *__restrict__ order,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ Cx)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const int64_t s1 = seg[e + 1];
    double acc = 0.0;
    for (int64_t s = seg[e]; s < s1; s += 4)
    {
        const int64_t cnt = s1 - s;
        unsigned t[4];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] = order[u < cnt ? s + u : s];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = x[t[u]];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (u < cnt) acc = acc + v[u];
    }
    __builtin_nontemporal_store(acc, Cx + e);
}

__global__ void k_set_i64(int64_t *p, int64_t v) { *p = v; }

__global__ void k_gather_lengths(int m_out, const int *__restrict__ map, const int *__restrict__ Ap,
                                 int *__restrict__ len)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m_out) len[r] = Ap[map[r] + 1] - Ap[map[r]];
}

__global__ void k_gather_sources(int m_out, const int *__restrict__ map,
                                 const int *__restrict__ Ap, const int64_t *__restrict__ Cp,
                                 int *__restrict__ src)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m_out) return;
    int a = Ap[map[r]];
    for (int64_t e = Cp[r]; e < Cp[r + 1]; e++) src[e] = a + (int) (e - Cp[r]);
}

__global__ void k_scale_rows(int64_t nnz, const int *__restrict__ rowidx,
                             const double *__restrict__ Ax, const double *__restrict__ d,
                             double *__restrict__ Cx)
{
    int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nnz) Cx[t] = Ax[t] * d[rowidx[t]];
}

__global__ void k_kron_spmv(int m, int n, int p, const int *__restrict__ Ap,
                            const int *__restrict__ Ai, const double *__restrict__ Ax,
                            const double *__restrict__ x, double *__restrict__ y)
{
    int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t) m * p) return;
    int b = (int) (g / m), i = (int) (g - (int64_t) b * m);
    const double *xb = x + (int64_t) b * n;
    double s = 0.0;
    for (int t = Ap[i]; t < Ap[i + 1]; t++) s = s + Ax[t] * xb[Ai[t]];
    y[g] = s;
}

__global__ void k_narrow(int64_t n, const int64_t *__restrict__ in, int *__restrict__ out)
{
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int) in[i];
}

__global__ void k_lower_count(int m, const int *__restrict__ Ap, const int *__restrict__ Ai,
                              int *__restrict__ cnt)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int c = 0;
    for (int t = Ap[r]; t < Ap[r + 1]; t++) c += Ai[t] <= r;
    cnt[r] = c;
}

__global__ void k_lower_write(int m, const int *__restrict__ Ap, const int *__restrict__ Ai,
                              const int64_t *__restrict__ off, int *__restrict__ rows,
                              int *__restrict__ cols, int *__restrict__ vmap)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int64_t o = off[r];
    for (int t = Ap[r]; t < Ap[r + 1]; t++)
        if (Ai[t] <= r)
        {
            rows[o] = r;
            cols[o] = Ai[t];
            vmap[o] = t;
            o++;
        }
}

template <typename K>
void coalesce_impl(CoalescePlan &P, int64_t n, const int *dorow, const int *docol, int m_out,
                   int n_out)
{
    const int cbits = bits_for(std::max(n_out, 1));
    const int rbits = bits_for(std::max(m_out, 1));
    const K cmask = (K) (((uint64_t) 1 << cbits) - 1);
    DBuf<K> keys(n);
    P.order.alloc(n);
    hipLaunchKernelGGL(k_make_keys<K>, dim3(grid_for(n)), dim3(256), 0, stream(), n, dorow,
                       docol, cbits, keys.p);
    iota_u32(P.order.p, n);
    if (sizeof(K) == 4) radix_sort_u32((uint32_t *) keys.p, P.order.p, n, rbits + cbits);
    else radix_sort_u64((uint64_t *) keys.p, P.order.p, n, rbits + cbits);
    DBuf<int> head(n);
    hipLaunchKernelGGL(k_heads<K>, dim3(grid_for(n)), dim3(256), 0, stream(), n, keys.p, head.p);
    DBuf<int64_t> hoff(n + 1);
    P.nnz = scan_i32_to_i64(head.p, hoff.p, n, true);
    P.Ci.alloc(P.nnz);
    P.seg.alloc(P.nnz + 1);
    DBuf<K> ekeys(P.nnz);
    hipLaunchKernelGGL(k_coalesce_entries<K>, dim3(grid_for(n)), dim3(256), 0, stream(), n,
                       keys.p, head.p, hoff.p, cmask, P.Ci.p, P.seg.p, ekeys.p);
    hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, stream(), P.seg.p + P.nnz, n);
    hipLaunchKernelGGL(k_bounds_from_sorted<K>, dim3(grid_for((int64_t) m_out + 1)), dim3(256),
                       0, stream(), ekeys.p, P.nnz, (int64_t) m_out, cbits, P.Cp.p,
                       (int *) nullptr);
    SDL_HIP(hipGetLastError());
    sync();
}

} // namespace

void expand_rows(int m, const int *dAp, int *drowidx, hipStream_t s)
{
    if (m <= 0) return;
    hipLaunchKernelGGL(k_expand_rows, dim3(grid_for(m)), dim3(256), 0, s, m, dAp, drowidx);
    SDL_HIP(hipGetLastError());
}

void transpose_structure(int m, int n, int nnz, const int *dAp, const int *dAi, int *dATp,
                         int *dATi, int *dperm)
{
    DBuf<unsigned> keys(nnz), vals(nnz);
    DBuf<int> rowidx(nnz);
    if (nnz > 0)
    {
        hipLaunchKernelGGL(k_copy_keys_u32, dim3(grid_for(nnz)), dim3(256), 0, stream(),
                           (int64_t) nnz, dAi, keys.p);
        iota_u32(vals.p, nnz);
        radix_sort_u32(keys.p, vals.p, nnz, bits_for(std::max(n, 1)));
        expand_rows(m, dAp, rowidx.p, stream());
        hipLaunchKernelGGL(k_transpose_finish, dim3(grid_for(nnz)), dim3(256), 0, stream(),
                           (int64_t) nnz, vals.p, rowidx.p, dATi, dperm);
    }
    hipLaunchKernelGGL(k_bounds_from_sorted<unsigned>, dim3(grid_for((int64_t) n + 1)),
                       dim3(256), 0, stream(), keys.p, (int64_t) nnz, (int64_t) n, 0,
                       (int64_t *) nullptr, dATp);
    SDL_HIP(hipGetLastError());
    sync();
}

void gather_values(int64_t n, const int *dperm, const double *din, double *dout, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_values, dim3(grid_for(n)), dim3(256), 0, s, n, dperm, din, dout);
    SDL_HIP(hipGetLastError());
}

void coalesce_plan(CoalescePlan &P, int64_t n_items, const int *dorow, const int *docol,
                   int m_out, int n_out)
{
    P.m_out = m_out;
    P.n_out = n_out;
    P.n_items = n_items;
    P.Cp.alloc((size_t) m_out + 1);
    if (n_items == 0)
    {
        P.nnz = 0;
        dzero(P.Cp.p, P.Cp.bytes());
        P.Ci.alloc(0);
        P.seg.alloc(1);
        dzero(P.seg.p, sizeof(int64_t));
        P.order.alloc(0);
        sync();
        return;
    }
    int total_bits = bits_for(std::max(n_out, 1)) + bits_for(std::max(m_out, 1));
    if (total_bits <= 32) coalesce_impl<unsigned>(P, n_items, dorow, docol, m_out, n_out);
    else coalesce_impl<unsigned long long>(P, n_items, dorow, docol, m_out, n_out);
}

void coalesce_fill(const CoalescePlan &P, const double *dx, double *dCx, hipStream_t s)
{
    if (P.nnz <= 0) return;
    hipLaunchKernelGGL(k_coalesce_fill, dim3(grid_for(P.nnz)), dim3(256), 0, s, P.nnz, P.seg.p,
                       P.order.p, dx, dCx);
    SDL_HIP(hipGetLastError());
}

int64_t row_gather_plan(int m_out, const int *dmap, const int *dAp, int64_t *dCp64)
{
    DBuf<int> len(m_out);
    if (m_out > 0)
        hipLaunchKernelGGL(k_gather_lengths, dim3(grid_for(m_out)), dim3(256), 0, stream(), m_out,
                           dmap, dAp, len.p);
    SDL_HIP(hipGetLastError());
    return scan_i32_to_i64(len.p, dCp64, m_out, true);
}

void row_gather_sources(int m_out, const int *dmap, const int *dAp, const int64_t *dCp64,
                        int64_t nnz, int *dsrc)
{
    (void) nnz;
    if (m_out <= 0) return;
    hipLaunchKernelGGL(k_gather_sources, dim3(grid_for(m_out)), dim3(256), 0, stream(), m_out,
                       dmap, dAp, dCp64, dsrc);
    SDL_HIP(hipGetLastError());
}

void scale_rows(int64_t nnz, const int *drowidx, const double *dAx, const double *dd,
                double *dCx, hipStream_t s)
{
    if (nnz <= 0) return;
    hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(nnz)), dim3(256), 0, s, nnz, drowidx, dAx, dd,
                       dCx);
    SDL_HIP(hipGetLastError());
}

void kron_spmv(int m, int n, const int *dAp, const int *dAi, const double *dAx,
               const double *dx, double *dy, int p, hipStream_t s)
{
    int64_t tot = (int64_t) m * p;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_kron_spmv, dim3(grid_for(tot)), dim3(256), 0, s, m, n, p, dAp, dAi, dAx,
                       dx, dy);
    SDL_HIP(hipGetLastError());
}

// Lower-triangular COO export (COO_matrix.c:48-90): entries with col <= row in
// CSR order, value_map = CSR position.
void lower_coo_structure(int m, int nnz, const int *dAp, const int *dAi, DBuf<int> &rows,
                         DBuf<int> &cols, DBuf<int> &vmap, int &count)
{
    (void) nnz;
    DBuf<int> cnt(m);
    DBuf<int64_t> off((size_t) m + 1);
    if (m > 0)
        hipLaunchKernelGGL(k_lower_count, dim3(grid_for(m)), dim3(256), 0, stream(), m, dAp, dAi,
                           cnt.p);
    int64_t tot = scan_i32_to_i64(cnt.p, off.p, m, true);
    count = (int) tot;
    rows.alloc(tot);
    cols.alloc(tot);
    vmap.alloc(tot);
    if (m > 0)
        hipLaunchKernelGGL(k_lower_write, dim3(grid_for(m)), dim3(256), 0, stream(), m, dAp, dAi,
                           off.p, rows.p, cols.p, vmap.p);
    SDL_HIP(hipGetLastError());
    sync();
}

void narrow_i64(int64_t n, const int64_t *din, int *dout, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(n)), dim3(256), 0, s, n, din, dout);
    SDL_HIP(hipGetLastError());
}

} // namespace sdl
