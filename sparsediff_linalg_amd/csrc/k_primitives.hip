// Device-wide primitives: exclusive scan (int32/int64 -> int64) and a stable
// LSD radix sort of (key, u32 value) pairs.  Both are alloc-phase tools: they
// allocate their scratch.  Wavefront-width (64) ballots drive the in-tile
// stable ranking of the sort, so no per-thread digit counters are needed.
#include "sdl_internal.h"

namespace sdl {

namespace {

constexpr int SC_BLOCK = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = SC_BLOCK * SC_ITEMS;

__device__ __forceinline__ long long wave_incl_scan(long long x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        long long y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ __forceinline__ long long block_excl_scan(long long x, long long *total)
{
    __shared__ long long wsum[SC_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long inc = wave_incl_scan(x);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    long long wpre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SC_BLOCK / 64; i++)
    {
        long long s = wsum[i];
        if (i < w) wpre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wpre + inc - x;
}

template <typename T>
__global__ void __launch_bounds__(SC_BLOCK) k_scan_reduce(const T *in, int64_t n, long long *bsum)
{
    int64_t base = (int64_t) blockIdx.x * SC_TILE + (int64_t) threadIdx.x * SC_ITEMS;
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++)
        if (base + i < n) s += (long long) in[base + i];
    long long tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <typename T>
__global__ void __launch_bounds__(SC_BLOCK)
    k_scan_apply(const T *in, int64_t n, const long long *boff, int64_t *out)
{
    int64_t base = (int64_t) blockIdx.x * SC_TILE + (int64_t) threadIdx.x * SC_ITEMS;
    long long v[SC_ITEMS];
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++)
    {
        v[i] = (base + i < n) ? (long long) in[base + i] : 0;
        s += v[i];
    }
    long long tot;
    long long pre = block_excl_scan(s, &tot) + (boff ? boff[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++)
    {
        if (base + i < n) out[base + i] = pre;
        pre += v[i];
    }
    if (base < n && base + SC_ITEMS >= n) out[n] = pre;  // thread holding the last element
}

template <typename T> int64_t scan_impl(const T *in, int64_t *out, int64_t n, bool want_total)
{
    if (n == 0)
    {
        dzero(out, sizeof(int64_t));
        return 0;
    }
    int64_t nb = (n + SC_TILE - 1) / SC_TILE;
    if (nb == 1)
    {
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(SC_BLOCK), 0, stream(), in, n,
                           (const long long *) nullptr, out);
    }
    else
    {
        DBuf<long long> bsum(nb);
        DBuf<int64_t> boff(nb + 1);
        hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned) nb), dim3(SC_BLOCK), 0, stream(),
                           in, n, bsum.p);
        scan_impl<long long>(bsum.p, (int64_t *) boff.p, nb, false);
        hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned) nb), dim3(SC_BLOCK), 0, stream(),
                           in, n, (const long long *) boff.p, out);
        SDL_HIP(hipGetLastError());
        if (!want_total)
        {
            // keep bsum/boff alive until the kernels that read them have run
            sync();
            return 0;
        }
    }
    SDL_HIP(hipGetLastError());
    int64_t total = 0;
    if (want_total) d2h(&total, out + n, sizeof(int64_t));
    else sync();
    return total;
}

// ------------------------------------------------------------ radix sort ----
constexpr int RS_BLOCK = 256;  // 4 wavefronts
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_ITEMS = 16;   // rounds of 64 consecutive items per wavefront
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_BINS = 256;

template <typename K>
__global__ void __launch_bounds__(RS_BLOCK)
    k_rs_hist(const K *keys, int64_t n, int shift, unsigned *hist, int ntiles)
{
    __shared__ unsigned cnt[RS_BINS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t) blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int it = 0; it < RS_ITEMS; it++)
    {
        int64_t idx = base + (int64_t) it * RS_BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&cnt[(unsigned) (keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t) threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

template <typename K>
__global__ void __launch_bounds__(RS_BLOCK)
    k_rs_scatter(const K *__restrict__ kin, const unsigned *__restrict__ vin,
                 K *__restrict__ kout, unsigned *__restrict__ vout, int64_t n, int shift,
                 const int64_t *__restrict__ offs, int ntiles)
{
    __shared__ unsigned wcnt[RS_WAVES][RS_BINS];
    __shared__ int64_t gbase[RS_BINS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < RS_WAVES; i++) wcnt[i][threadIdx.x] = 0;
    gbase[threadIdx.x] = offs[(int64_t) threadIdx.x * ntiles + blockIdx.x];
    __syncthreads();

    const int64_t wbase = (int64_t) blockIdx.x * RS_TILE + (int64_t) w * (RS_ITEMS * 64);
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    K key[RS_ITEMS];
    unsigned val[RS_ITEMS];
    unsigned rank[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++)
    {
        int64_t idx = wbase + (int64_t) r * 64 + lane;
        bool valid = idx < n;
        key[r] = valid ? kin[idx] : (K) 0;
        val[r] = valid ? vin[idx] : 0u;
        unsigned d = (unsigned) (key[r] >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++)
        {
            unsigned long long bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        unsigned pre = (unsigned) __popcll(peers & below);
        // All lanes of the wavefront read before the group leader writes; LDS
        // operations of one wavefront execute in issue order.
        unsigned before = valid ? wcnt[w][d] : 0u;
        if (valid && pre == 0) wcnt[w][d] = before + (unsigned) __popcll(peers);
        rank[r] = before + pre;
    }
    __syncthreads();
    {
        unsigned run = 0;
#pragma unroll
        for (int i = 0; i < RS_WAVES; i++)
        {
            unsigned c = wcnt[i][threadIdx.x];
            wcnt[i][threadIdx.x] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++)
    {
        int64_t idx = wbase + (int64_t) r * 64 + lane;
        if (idx < n)
        {
            unsigned d = (unsigned) (key[r] >> shift) & 255u;
            int64_t pos = gbase[d] + wcnt[w][d] + rank[r];
            kout[pos] = key[r];
            vout[pos] = val[r];
        }
    }
}

template <typename K> void radix_impl(K *keys, unsigned *vals, int64_t n, int bits)
{
    if (n <= 1 || bits <= 0) return;
    const int passes = (bits + 7) / 8;
    const int64_t ntiles64 = (n + RS_TILE - 1) / RS_TILE;
    SDL_REQUIRE(ntiles64 < (1ll << 31), "radix sort: too many tiles");
    const int ntiles = (int) ntiles64;
    DBuf<K> ktmp(n);
    DBuf<unsigned> vtmp(n);
    DBuf<unsigned> hist((size_t) RS_BINS * ntiles);
    DBuf<int64_t> offs((size_t) RS_BINS * ntiles + 1);
    K *ka = keys, *kb = ktmp.p;
    unsigned *va = vals, *vb = vtmp.p;
    for (int p = 0; p < passes; p++)
    {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_rs_hist<K>, dim3(ntiles), dim3(RS_BLOCK), 0, stream(), ka, n,
                           shift, hist.p, ntiles);
        scan_impl<int>((const int *) hist.p, offs.p, (int64_t) RS_BINS * ntiles, false);
        hipLaunchKernelGGL(k_rs_scatter<K>, dim3(ntiles), dim3(RS_BLOCK), 0, stream(), ka,
                           va, kb, vb, n, shift, offs.p, ntiles);
        SDL_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys)
    {
        d2d(keys, ka, (size_t) n * sizeof(K));
        d2d(vals, va, (size_t) n * sizeof(unsigned));
    }
    sync();  // scratch is freed on return
}

__global__ void k_iota(unsigned *v, int64_t n)
{
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (unsigned) i;
}

} // namespace

int64_t scan_i32_to_i64(const int *in, int64_t *out, int64_t n, bool want_total)
{
    return scan_impl<int>(in, out, n, want_total);
}

int64_t scan_i64(const int64_t *in, int64_t *out, int64_t n, bool want_total)
{
    return scan_impl<long long>((const long long *) in, out, n, want_total);
}

void radix_sort_u32(uint32_t *keys, uint32_t *vals, int64_t n, int bits)
{
    radix_impl<unsigned>(keys, vals, n, bits);
}

void radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bits)
{
    radix_impl<unsigned long long>((unsigned long long *) keys, vals, n, bits);
}

void iota_u32(uint32_t *v, int64_t n)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, stream(), v,
                       n);
    SDL_HIP(hipGetLastError());
}

} // namespace sdl
