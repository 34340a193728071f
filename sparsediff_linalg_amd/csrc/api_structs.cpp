// Drop-in raw-struct API: include/sparsediff_linalg/{CSR,CSC,COO}_matrix.h.
// Structs and their arrays live on the host (the ABI); every alloc computes
// structure on the GPU and binds the device state its fill needs to the
// output; every fill uploads the operands' values, runs the GPU kernels and
// writes the output's host values.
#include "api_common.h"
#include "mlp.h"

#include <climits>
#include <cstring>
#include <malloc.h>
#include <mutex>
#include <unordered_map>
#include <vector>

extern "C" {
#pragma GCC visibility push(default)
size_t g_allocated_bytes = 0;
size_t g_peak_bytes = 0;
#pragma GCC visibility pop
}

namespace sdl {

namespace {
std::mutex g_reg_mu;
std::unordered_map<const void *, std::unique_ptr<Attached>> &registry()
{
    static std::unordered_map<const void *, std::unique_ptr<Attached>> r;
    return r;
}
} // namespace

[[noreturn]] void api_exit(const char *fn, const char *msg)
{
    std::fprintf(stderr, "Error in %s: %s\n", fn, msg);
    std::fflush(stderr);
    std::exit(1);
}

int checked_nnz(const char *fn, int64_t nnz)
{
    if (nnz > INT_MAX) api_exit(fn, "output nnz exceeds INT_MAX.");
    return (int) nnz;
}

void attach(const void *owner, std::unique_ptr<Attached> st)
{
    std::lock_guard<std::mutex> g(g_reg_mu);
    registry()[owner] = std::move(st);
}

Attached *attached(const void *owner)
{
    std::lock_guard<std::mutex> g(g_reg_mu);
    auto it = registry().find(owner);
    return it == registry().end() ? nullptr : it->second.get();
}

void detach(const void *owner)
{
    std::unique_ptr<Attached> victim;
    {
        std::lock_guard<std::mutex> g(g_reg_mu);
        auto it = registry().find(owner);
        if (it == registry().end()) return;
        victim = std::move(it->second);
        registry().erase(it);
    }
    sync();
}

// The host-byte counters of tracked_alloc.h (tracked_alloc.h:40-52 of the
// reference): every ABI array the library allocates or frees is counted.
std::mutex g_track_mu;

void count_alloc(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> g(g_track_mu);
    g_allocated_bytes += malloc_usable_size(p);
    if (g_allocated_bytes > g_peak_bytes) g_peak_bytes = g_allocated_bytes;
}

void *host_malloc(size_t bytes)
{
    void *p = std::malloc(bytes > 0 ? bytes : 1);
    count_alloc(p);
    return p;
}

void host_free(void *p)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> g(g_track_mu);
        g_allocated_bytes -= malloc_usable_size(p);
    }
    std::free(p);
}

int *host_ints(int64_t n, bool zero)
{
    size_t bytes = (size_t) (n > 0 ? n : 1) * sizeof(int);
    void *p = zero ? std::calloc(1, bytes) : std::malloc(bytes);
    count_alloc(p);
    return (int *) p;
}

double *host_doubles(int64_t n)
{
    return (double *) host_malloc((size_t) (n > 0 ? n : 1) * sizeof(double));
}

namespace {

// ---- bound states -------------------------------------------------------
struct TransposeState : Attached
{
    int src_m = 0, src_n = 0, nnz = 0;
    uint64_t fp = 0;
    DBuf<int> perm;
    DBuf<double> xin, xout;
};

struct ProductState : Attached
{
    int Am = 0, An = 0, nnzA = 0, Lm = 0, nnzL = 0;
    uint64_t fpA = 0, fpL = 0;
    DBuf<int> Ap, Ai, Lp, Li;
    ProductPlan P;
    DBuf<double> Ax, Lx, d, Cx;
};

struct CoalesceState : Attached
{
    int64_t nnz_in = 0;
    uint64_t fp = 0;
    CoalescePlan P;
    DBuf<double> x, Cx;
};

struct LowerCOOState : Attached
{
    int n_src = 0;
    DBuf<int> vmap;
    DBuf<double> vals, x;
};

// Fingerprint of a compressed structure, taken at alloc and checked by every
// fill: the two dimensions, all np pointers and up to 4096 evenly spaced
// entries of each index array, mixed as 32-bit words (FNV-1a over words).  A
// fill handed another structure of the same nnz exits rather than gathering
// through the alloc's permutation; O(np) per fill, so a change confined to
// unsampled indices of an unchanged pointer array goes unseen.
uint64_t fingerprint(int d0, int d1, int np, const int *p, int64_t nnz, const int *i,
                     const int *j = nullptr)
{
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&h](uint32_t w) {
        h ^= w;
        h *= 0x100000001b3ull;
    };
    mix((uint32_t) d0);
    mix((uint32_t) d1);
    mix((uint32_t) nnz);
    mix((uint32_t) (nnz >> 32));
    for (int k = 0; k < np; k++) mix((uint32_t) p[k]);
    const int64_t S = std::min<int64_t>(nnz, 4096);
    for (int64_t k = 0; k < S; k++)
    {
        const int64_t at = k * nnz / S;
        mix((uint32_t) i[at]);
        if (j) mix((uint32_t) j[at]);
    }
    return h;
}

CSR_matrix *alloc_csr_struct(int m, int n, int nnz)
{
    CSR_matrix *C = (CSR_matrix *) host_malloc(sizeof(CSR_matrix));
    C->p = host_ints((int64_t) m + 1, true);
    C->i = host_ints(nnz, true);
    C->x = host_doubles(nnz);
    C->m = m;
    C->n = n;
    C->nnz = nnz;
    return C;
}

// Transpose structure of an m x n compressed matrix into a new (n x m) struct
// of the same layout; binds the permutation for the fill.
template <typename Out>
Out *transpose_alloc_impl(const char *fn, int m, int n, int nnz, const int *p, const int *i)
{
    auto st = std::make_unique<TransposeState>();
    st->src_m = m;
    st->src_n = n;
    st->nnz = nnz;
    st->fp = fingerprint(m, n, m + 1, p, nnz, i);
    DBuf<int> dp = upload(p, (int64_t) m + 1), di = upload(i, nnz);
    DBuf<int> tp((size_t) n + 1), ti(nnz);
    st->perm.alloc(nnz);
    transpose_structure(m, n, nnz, dp.p, di.p, tp.p, ti.p, st->perm.p);
    st->xin.alloc(nnz);
    st->xout.alloc(nnz);
    Out *C = (Out *) alloc_csr_struct(0, 0, nnz);
    host_free(C->p);
    C->p = host_ints((int64_t) n + 1, false);
    C->m = m;
    C->n = n;  // fixed up by the caller for the CSR (n x m) case
    d2h(C->p, tp.p, ((size_t) n + 1) * sizeof(int));
    d2h(C->i, ti.p, (size_t) nnz * sizeof(int));
    (void) fn;
    attach(C, std::move(st));
    return C;
}

void transpose_fill_impl(const char *fn, const void *C, int m, int n, int nnz, const int *p,
                         const int *i, const double *xin, double *xout)
{
    TransposeState *st = attached_as<TransposeState>(C, fn);
    if (st->nnz != nnz || st->src_m != m || st->src_n != n ||
        st->fp != fingerprint(m, n, m + 1, p, nnz, i))
        api_exit(fn, "operand does not match the alloc");
    h2d(st->xin.p, xin, (size_t) nnz * sizeof(double));
    gather_values(nnz, st->perm.p, st->xin.p, st->xout.p, stream());
    d2h(xout, st->xout.p, (size_t) nnz * sizeof(double));
}

// C = kron(I,L) D A with L given as host compressed rows (CSC columns).
CSR_matrix *product_alloc_impl(const char *fn, int Lm, const int *Lp, const int *Li, int Am,
                               int An, const int *Ap, const int *Ai, bool allow_pull)
{
    auto st = std::make_unique<ProductState>();
    st->Am = Am;
    st->An = An;
    st->nnzA = Ap[Am];
    st->Lm = Lm;
    st->nnzL = Lp[Lm];
    st->Ap = upload(Ap, (int64_t) Am + 1);
    st->Ai = upload(Ai, st->nnzA);
    st->Lp = upload(Lp, (int64_t) Lm + 1);
    st->Li = upload(Li, st->nnzL);
    st->fpA = fingerprint(Am, An, Am + 1, Ap, st->nnzA, Ai);
    st->fpL = fingerprint(Lm, Am, Lm + 1, Lp, st->nnzL, Li);
    LView L;
    L.Lm = Lm;
    L.Lp = st->Lp.p;
    L.Li = st->Li.p;
    AView A;
    A.Am = Am;
    A.An = An;
    A.Ap = st->Ap.p;
    A.Ai = st->Ai.p;
    product_plan(st->P, L, A, g_row_predictor ? SymMode::Predicted : SymMode::Exact, g_row_predictor,
                 allow_pull);
    int nnz = checked_nnz(fn, st->P.nnz);
    CSR_matrix *C = alloc_csr_struct(Lm, An, nnz);
    DBuf<int> p32((size_t) Lm + 1);
    narrow_i64((int64_t) Lm + 1, st->P.Cp.p, p32.p, stream());
    d2h(C->p, p32.p, ((size_t) Lm + 1) * sizeof(int));
    d2h(C->i, st->P.Ci.p, (size_t) nnz * sizeof(int));
    st->Ax.alloc(st->nnzA);
    st->Lx.alloc(st->nnzL);
    st->d.alloc(Am);
    st->Cx.alloc(nnz);
    attach(C, std::move(st));
    return C;
}

void product_fill_impl(const char *fn, const int *Lp, const int *Li, const double *Lx,
                       int nnzL, const double *d, const int *Ap, const int *Ai,
                       const double *Ax, int nnzA, CSR_matrix *C)
{
    ProductState *st = attached_as<ProductState>(C, fn);
    if (st->nnzA != nnzA || st->nnzL != nnzL || C->nnz != st->P.nnz ||
        st->fpA != fingerprint(st->Am, st->An, st->Am + 1, Ap, nnzA, Ai) ||
        st->fpL != fingerprint(st->Lm, st->Am, st->Lm + 1, Lp, nnzL, Li))
        api_exit(fn, "operands do not match the alloc");
    h2d(st->Ax.p, Ax, (size_t) nnzA * sizeof(double));
    h2d(st->Lx.p, Lx, (size_t) nnzL * sizeof(double));
    if (d) h2d(st->d.p, d, (size_t) st->Am * sizeof(double));
    product_numeric(st->P, st->Lx.p, d ? st->d.p : nullptr, st->Ax.p, st->Cx.p, stream());
    d2h(C->x, st->Cx.p, (size_t) C->nnz * sizeof(double));
}

} // namespace
} // namespace sdl

using namespace sdl;

extern "C" {
#pragma GCC visibility push(default)

// ------------------------------------------------- tracked_alloc.h ----
void sp_memory_stats(size_t *live_bytes, size_t *peak_bytes)
{
    std::lock_guard<std::mutex> g(g_track_mu);
    if (live_bytes) *live_bytes = g_allocated_bytes;
    if (peak_bytes) *peak_bytes = g_peak_bytes;
}

void sp_memory_reset_peak(void)
{
    std::lock_guard<std::mutex> g(g_track_mu);
    g_peak_bytes = g_allocated_bytes;
}

// ------------------------------------------------------------- CSR ----
CSR_matrix *new_CSR_matrix(int m, int n, int nnz) { return alloc_csr_struct(m, n, nnz); }

CSR_matrix *new_csr(const CSR_matrix *A)
{
    CSR_matrix *C = alloc_csr_struct(A->m, A->n, A->nnz);
    std::memcpy(C->p, A->p, ((size_t) A->m + 1) * sizeof(int));
    std::memcpy(C->i, A->i, (size_t) A->nnz * sizeof(int));
    std::memcpy(C->x, A->x, (size_t) A->nnz * sizeof(double));
    return C;
}

void free_CSR_matrix(CSR_matrix *matrix)
{
    if (!matrix) return;
    detach(matrix);
    host_free(matrix->p);
    host_free(matrix->i);
    host_free(matrix->x);
    host_free(matrix);
}

void copy_CSR_matrix(const CSR_matrix *A, CSR_matrix *C)
{
    C->m = A->m;
    C->n = A->n;
    C->nnz = A->nnz;
    std::memcpy(C->p, A->p, ((size_t) A->m + 1) * sizeof(int));
    std::memcpy(C->i, A->i, (size_t) A->nnz * sizeof(int));
    std::memcpy(C->x, A->x, (size_t) A->nnz * sizeof(double));
}

void CSR_trim(CSR_matrix *A)
{
    int nnz = A->p[A->m];
    A->nnz = nnz;
    if (nnz == 0) return;
    A->i = (int *) std::realloc(A->i, (size_t) nnz * sizeof(int));
    A->x = (double *) std::realloc(A->x, (size_t) nnz * sizeof(double));
}

CSR_matrix *AT_alloc(const CSR_matrix *A, int *iwork)
{
    (void) iwork;
    CSR_matrix *C = transpose_alloc_impl<CSR_matrix>("AT_alloc", A->m, A->n, A->nnz, A->p, A->i);
    C->m = A->n;
    C->n = A->m;
    return C;
}

void AT_fill_values(const CSR_matrix *A, CSR_matrix *AT, int *iwork)
{
    (void) iwork;
    transpose_fill_impl("AT_fill_values", AT, A->m, A->n, A->nnz, A->p, A->i, A->x, AT->x);
}

CSR_matrix *transpose(const CSR_matrix *A, int *iwork)
{
    CSR_matrix *AT = AT_alloc(A, iwork);
    AT_fill_values(A, AT, iwork);
    return AT;
}

// ------------------------------------------------------------- CSC ----
CSC_matrix *new_CSC_matrix(int m, int n, int nnz)
{
    CSC_matrix *C = (CSC_matrix *) host_malloc(sizeof(CSC_matrix));
    C->p = host_ints((int64_t) n + 1, false);
    C->i = host_ints(nnz, false);
    C->x = host_doubles(nnz);
    C->m = m;
    C->n = n;
    C->nnz = nnz;
    return C;
}

void free_CSC_matrix(CSC_matrix *matrix)
{
    if (!matrix) return;
    detach(matrix);
    host_free(matrix->p);
    host_free(matrix->i);
    host_free(matrix->x);
    host_free(matrix);
}

CSC_matrix *csr_to_csc_alloc(const CSR_matrix *A, int *iwork)
{
    (void) iwork;
    CSC_matrix *C =
        transpose_alloc_impl<CSC_matrix>("csr_to_csc_alloc", A->m, A->n, A->nnz, A->p, A->i);
    C->m = A->m;
    C->n = A->n;
    return C;
}

void csr_to_csc_fill_values(const CSR_matrix *A, CSC_matrix *C, int *iwork)
{
    (void) iwork;
    transpose_fill_impl("csr_to_csc_fill_values", C, A->m, A->n, A->nnz, A->p, A->i, A->x,
                        C->x);
}

CSR_matrix *csc_to_csr_alloc(const CSC_matrix *A, int *iwork)
{
    (void) iwork;
    // A's columns are the rows of the compressed source (n x m view)
    CSR_matrix *C =
        transpose_alloc_impl<CSR_matrix>("csc_to_csr_alloc", A->n, A->m, A->nnz, A->p, A->i);
    C->m = A->m;
    C->n = A->n;
    return C;
}

void csc_to_csr_fill_values(const CSC_matrix *A, CSR_matrix *C, int *iwork)
{
    (void) iwork;
    transpose_fill_impl("csc_to_csr_fill_values", C, A->n, A->m, A->nnz, A->p, A->i, A->x,
                        C->x);
}

CSR_matrix *BTA_alloc(const CSR_matrix *A, const CSC_matrix *B)
{
    if (A->m != B->m) api_exit("BTA_alloc", "A and B must have the same number of rows.");
    // the raw-struct fill is the reference's accumulator fill (CSC_matrix.c:198-203);
    // both GPU schedules give identical values, so the faster one is used.
    return product_alloc_impl("BTA_alloc", B->n, B->p, B->i, A->m, A->n, A->p, A->i, true);
}

void BTDA_fill_values(const CSR_matrix *A, const CSC_matrix *B, const double *d, double *acc,
                      CSR_matrix *C)
{
    (void) acc;
    product_fill_impl("BTDA_fill_values", B->p, B->i, B->x, B->nnz, d, A->p, A->i, A->x,
                      A->nnz, C);
}

CSR_matrix *ATA_alloc(const CSC_matrix *A_csc, const CSR_matrix *A)
{
    if (A_csc->m != A->m || A_csc->n != A->n) api_exit("ATA_alloc", "shape mismatch.");
    return product_alloc_impl("ATA_alloc", A_csc->n, A_csc->p, A_csc->i, A->m, A->n, A->p,
                              A->i, true);
}

void ATDA_fill_values(const CSC_matrix *A_csc, const double *d, const CSR_matrix *A,
                      double *acc, CSR_matrix *C)
{
    (void) acc;
    product_fill_impl("ATDA_fill_values", A_csc->p, A_csc->i, A_csc->x, A_csc->nnz, d, A->p,
                      A->i, A->x, A->nnz, C);
}

// ------------------------------------------------------------- COO ----
COO_matrix *new_COO_matrix(const CSR_matrix *A)
{
    COO_matrix *coo = (COO_matrix *) host_malloc(sizeof(COO_matrix));
    coo->m = A->m;
    coo->n = A->n;
    coo->nnz = A->nnz;
    coo->rows = host_ints(A->nnz, false);
    coo->cols = host_ints(A->nnz, false);
    coo->x = host_doubles(A->nnz);
    coo->value_map = nullptr;
    DBuf<int> dp = upload(A->p, (int64_t) A->m + 1);
    DBuf<int> rows(A->nnz);
    expand_rows(A->m, dp.p, rows.p, stream());
    d2h(coo->rows, rows.p, (size_t) A->nnz * sizeof(int));
    std::memcpy(coo->cols, A->i, (size_t) A->nnz * sizeof(int));
    std::memcpy(coo->x, A->x, (size_t) A->nnz * sizeof(double));
    return coo;
}

void free_COO_matrix(COO_matrix *matrix)
{
    if (!matrix) return;
    detach(matrix);
    host_free(matrix->rows);
    host_free(matrix->cols);
    host_free(matrix->x);
    host_free(matrix->value_map);
    host_free(matrix);
}

#pragma GCC visibility pop
} // extern "C"

namespace sdl {
void lower_coo_structure(int m, int nnz, const int *dAp, const int *dAi, DBuf<int> &rows,
                         DBuf<int> &cols, DBuf<int> &vmap, int &count);
}

extern "C" {
#pragma GCC visibility push(default)

COO_matrix *new_COO_matrix_lower_triangular(const CSR_matrix *A)
{
    DBuf<int> dp = upload(A->p, (int64_t) A->m + 1), di = upload(A->i, A->nnz);
    DBuf<int> rows, cols;
    auto st = std::make_unique<LowerCOOState>();
    int count = 0;
    lower_coo_structure(A->m, A->nnz, dp.p, di.p, rows, cols, st->vmap, count);
    COO_matrix *coo = (COO_matrix *) host_malloc(sizeof(COO_matrix));
    coo->m = A->m;
    coo->n = A->n;
    coo->nnz = count;
    coo->rows = host_ints(count, false);
    coo->cols = host_ints(count, false);
    coo->x = host_doubles(count);
    coo->value_map = host_ints(count, false);
    d2h(coo->rows, rows.p, (size_t) count * sizeof(int));
    d2h(coo->cols, cols.p, (size_t) count * sizeof(int));
    d2h(coo->value_map, st->vmap.p, (size_t) count * sizeof(int));
    st->n_src = A->nnz;
    st->vals.alloc(A->nnz);
    st->x.alloc(count);
    h2d(st->vals.p, A->x, (size_t) A->nnz * sizeof(double));
    gather_values(count, st->vmap.p, st->vals.p, st->x.p, stream());
    d2h(coo->x, st->x.p, (size_t) count * sizeof(double));
    attach(coo, std::move(st));
    return coo;
}

void refresh_lower_triangular_coo(COO_matrix *coo, const double *vals)
{
    LowerCOOState *st = attached_as<LowerCOOState>(coo, "refresh_lower_triangular_coo");
    h2d(st->vals.p, vals, (size_t) st->n_src * sizeof(double));
    gather_values(coo->nnz, st->vmap.p, st->vals.p, st->x.p, stream());
    d2h(coo->x, st->x.p, (size_t) coo->nnz * sizeof(double));
}

CSR_matrix *COO_coalesce_to_csr_alloc(const COO_matrix *A)
{
    auto st = std::make_unique<CoalesceState>();
    st->nnz_in = A->nnz;
    st->fp = fingerprint(A->m, A->n, 0, nullptr, A->nnz, A->rows, A->cols);
    DBuf<int> r = upload(A->rows, A->nnz), c = upload(A->cols, A->nnz);
    coalesce_plan(st->P, A->nnz, r.p, c.p, A->m, A->n);
    int nnz = checked_nnz("COO_coalesce_to_csr_alloc", st->P.nnz);
    CSR_matrix *C = alloc_csr_struct(A->m, A->n, nnz);
    DBuf<int> p32((size_t) A->m + 1);
    narrow_i64((int64_t) A->m + 1, st->P.Cp.p, p32.p, stream());
    d2h(C->p, p32.p, ((size_t) A->m + 1) * sizeof(int));
    d2h(C->i, st->P.Ci.p, (size_t) nnz * sizeof(int));
    st->x.alloc(A->nnz);
    st->Cx.alloc(nnz);
    attach(C, std::move(st));
    return C;
}

void COO_coalesce_to_csr_fill_values(const COO_matrix *A, CSR_matrix *C)
{
    CoalesceState *st = attached_as<CoalesceState>(C, "COO_coalesce_to_csr_fill_values");
    if (st->nnz_in != A->nnz ||
        st->fp != fingerprint(A->m, A->n, 0, nullptr, A->nnz, A->rows, A->cols))
        api_exit("COO_coalesce_to_csr_fill_values", "COO does not match the alloc");
    h2d(st->x.p, A->x, (size_t) A->nnz * sizeof(double));
    coalesce_fill(st->P, st->x.p, st->Cx.p, stream());
    d2h(C->x, st->Cx.p, (size_t) C->nnz * sizeof(double));
}

#pragma GCC visibility pop
} // extern "C"
