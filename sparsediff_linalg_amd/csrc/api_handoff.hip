// Device-resident solver hand-off: the lower-triangular COO export of a
// product output that lives in HBM (reference: COO_matrix.c:48-98,
// new_COO_matrix_lower_triangular / refresh_lower_triangular_coo).  The
// reference exports entries with col <= row in CSR order with value_map = the
// CSR position of each, and refreshes x by gathering through value_map.  With
// sorted columns (every product plan's output) the exported entries of row r
// are the first cnt_r entries of the row, so the export is a row-prefix
// selection and the refresh a segmented copy: no map is read per refresh.
#include "sdl_internal.h"
#include "../../include/sdl_device.h"

#include <algorithm>
#include <new>

namespace sdl {

namespace {

unsigned lc_grid(int64_t n) { return (unsigned) std::max<int64_t>(1, (n + 255) / 256); }

// cnt[r] = number of entries of row r with column <= r (binary search in the
// sorted row).
__global__ void k_lower_prefix(int m, const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
                               int *__restrict__ cnt)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int64_t lo = Cp[r], hi = Cp[r + 1];
    const int64_t b = lo;
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) >> 1;
        if (Ci[mid] <= r) lo = mid + 1;
        else hi = mid;
    }
    cnt[r] = (int) (lo - b);
}

// Entry o of the export: row r with off[r] <= o < off[r+1] (binary search over
// the offsets), CSR position Cp[r] + (o - off[r]).
__global__ void k_lower_entries(int64_t total, int m, const int64_t *__restrict__ off,
                                const int64_t *__restrict__ Cp, const int *__restrict__ Ci,
                                int *__restrict__ rows, int *__restrict__ cols,
                                long long *__restrict__ vmap)
{
    int64_t o = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= total) return;
    int lo = 0, hi = m - 1;  // largest r with off[r] <= o
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= o) lo = mid;
        else hi = mid - 1;
    }
    const int64_t t = Cp[lo] + (o - off[lo]);
    rows[o] = lo;
    cols[o] = Ci[t];
    vmap[o] = t;
}

// x[off[r] + i] = Cx[Cp[r] + i] for i < off[r+1] - off[r]: LG lanes per row
// (four rows per wave), LU entries per lane per pass; every load of a pass is
// issued unconditionally (indices past the row's end re-read its last entry,
// whose copy is dropped), so a lane keeps LU loads in flight.  Reads and writes
// are contiguous runs of LG values.  (Measured slower: one wave per row, whose
// second pass leaves most lanes idle on ~65-entry rows; a wave per 256 export
// entries with a per-lane row walk, whose data-dependent walk serialises the
// loads; non-temporal stores, 19 % slower: a row's segment shares its first and
// last 128-byte lines with its neighbours', and ordinary stores let L2 merge
// those partial lines before they reach HBM; 32 lanes x 4 or 16 lanes x 4
// entries, 6 % slower.)
constexpr int LG = 16, LU = 8;
__global__ __launch_bounds__(256) void k_lower_refresh(int m, const int64_t *__restrict__ off,
                                                      const int64_t *__restrict__ Cp,
                                                      const double *__restrict__ Cx,
                                                      double *__restrict__ x)
{
    const int64_t r = ((int64_t) blockIdx.x * 256 + threadIdx.x) / LG;
    if (r >= m) return;  // no cross-lane operations below
    const int gl = threadIdx.x % LG;
    const int64_t d0 = off[r], len = off[r + 1] - d0, s0 = Cp[r];
    for (int64_t b = 0; b < len; b += LG * LU)
    {
        double v[LU];
#pragma unroll
        for (int u = 0; u < LU; u++)
        {
            const int64_t i = b + gl + LG * u;
            v[u] = Cx[s0 + (i < len ? i : len - 1)];
        }
#pragma unroll
        for (int u = 0; u < LU; u++)
        {
            const int64_t i = b + gl + LG * u;
            if (i < len) x[d0 + i] = v[u];
        }
    }
}

} // namespace

} // namespace sdl

struct sdl_lower_coo
{
    int m = 0;
    long long nnz = 0;
    sdl::DBuf<int64_t> Cp;      // copy of the source row pointers
    sdl::DBuf<int64_t> off;     // m + 1 export offsets
    sdl::DBuf<int> rows, cols;
    sdl::DBuf<long long> vmap;
};

extern "C" {
#pragma GCC visibility push(default)

int sdl_lower_coo_create(int m, const long long *dCp, const int *dCi, sdl_lower_coo **out)
{
    using namespace sdl;
    if (!out || m < 0 || (m > 0 && (!dCp || !dCi)))
    {
        set_error("sdl_lower_coo_create: invalid arguments");
        return 1;
    }
    sdl_lower_coo *h = new (std::nothrow) sdl_lower_coo();
    if (!h) return 1;
    h->m = m;
    h->Cp.alloc((size_t) m + 1);
    h->off.alloc((size_t) m + 1);
    d2d(h->Cp.p, dCp, ((size_t) m + 1) * sizeof(int64_t));
    DBuf<int> cnt(m);
    if (m > 0)
        hipLaunchKernelGGL(k_lower_prefix, dim3(lc_grid(m)), dim3(256), 0, stream(), m, h->Cp.p,
                           dCi, cnt.p);
    SDL_HIP(hipGetLastError());
    const int64_t total = scan_i32_to_i64(cnt.p, h->off.p, m, true);
    h->nnz = total;
    h->rows.alloc(total);
    h->cols.alloc(total);
    h->vmap.alloc(total);
    if (total > 0)
        hipLaunchKernelGGL(k_lower_entries, dim3(lc_grid(total)), dim3(256), 0, stream(), total, m,
                           h->off.p, h->Cp.p, dCi, h->rows.p, h->cols.p, h->vmap.p);
    SDL_HIP(hipGetLastError());
    sdl::sync();
    *out = h;
    return 0;
}

long long sdl_lower_coo_nnz(const sdl_lower_coo *h) { return h ? h->nnz : -1; }
const int *sdl_lower_coo_rows(const sdl_lower_coo *h) { return h ? h->rows.p : nullptr; }
const int *sdl_lower_coo_cols(const sdl_lower_coo *h) { return h ? h->cols.p : nullptr; }
const long long *sdl_lower_coo_value_map(const sdl_lower_coo *h)
{
    return h ? h->vmap.p : nullptr;
}

int sdl_lower_coo_refresh(const sdl_lower_coo *h, const double *dCx, double *dx, void *stream)
{
    using namespace sdl;
    if (!h || (h->nnz > 0 && (!dCx || !dx)))
    {
        set_error("sdl_lower_coo_refresh: invalid arguments");
        return 1;
    }
    if (h->m > 0 && h->nnz > 0)
        hipLaunchKernelGGL(k_lower_refresh,
                           dim3((unsigned) (((int64_t) h->m * LG + 255) / 256)), dim3(256), 0,
                           stream ? (hipStream_t) stream : sdl::stream(), h->m, h->off.p, h->Cp.p,
                           dCx, dx);
    SDL_HIP(hipGetLastError());
    return 0;
}

void sdl_lower_coo_destroy(sdl_lower_coo *h)
{
    if (!h) return;
    sdl::sync();
    delete h;
}

#pragma GCC visibility pop
} // extern "C"
