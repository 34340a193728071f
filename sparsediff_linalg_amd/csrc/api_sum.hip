// Drop-in matrix sums (reference: include/sparsediff_linalg/matrix_sum.h,
// src/sparse/matrix_sum.c, src/sparse/CSR_sum.c).  The alloc concatenates the
// operands' entries (operand 0 first, each in CSR order) as items, coalesces
// them on the GPU into the union structure and keeps the plan; the fill sums
// each output entry's items in that order from +0.0, one thread per entry, so
// C[r, c] = ((+0.0 + x_0) + x_1) + ... over the operands holding (r, c) -- the
// reference's accumulator order (CSR_sum.c:150-157 applied operand by
// operand, matrix_sum.c:142-155; the merge fills CSR_sum.c:84-148 give the
// same sums for canonical operands).
#include "api_matrix.h"
#include "api_pd.h"
#include "sparse_ops.h"

#include <climits>
#include <cstdlib>
#include <cstring>

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
__global__ void k_sum_map(int64_t nnz, const int64_t *__restrict__ seg,
                          const unsigned *__restrict__ order, int *__restrict__ map)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    for (int64_t s = seg[e]; s < seg[e + 1]; s++) map[order[s]] = (int) e;
}

// dup[k] = 1 when two items of operand k land in one entry (repeated (row,
// col) inside operand k).  Items of an entry are in ascending order, so
// repeats of one operand are neighbours.
__global__ void k_sum_dups(int64_t nnz, const int64_t *__restrict__ seg,
                           const unsigned *__restrict__ order, const int64_t *__restrict__ off,
                           int nops, int *__restrict__ dup)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    for (int64_t s = seg[e] + 1; s < seg[e + 1]; s++)
    {
        int k = op_of(off, nops, order[s]);
        if (k == op_of(off, nops, order[s - 1])) dup[k] = 1;
    }
}

// Cx[e] = ((+0.0 + v_1) + v_2) + ... over entry e's items in item order, with
// v = x_k[t - off[k]] or, SCALED, the rounded product d[k*m + row[t]] * x.
template <bool SCALED>
__global__ void k_sum_fill(int64_t nnz, const int64_t *__restrict__ seg,
                           const unsigned *__restrict__ order, const int64_t *__restrict__ off,
                           int nops, const double *const *__restrict__ xs,
                           const int *__restrict__ row, const double *__restrict__ d, int m,
                           double *__restrict__ Cx)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double acc = 0.0;
    for (int64_t s = seg[e]; s < seg[e + 1]; s++)
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
__global__ void k_sum_add_ordered(int64_t nnz, const int64_t *__restrict__ seg,
                                  const unsigned *__restrict__ order, int64_t lo, int64_t hi,
                                  const double *__restrict__ x, double *__restrict__ Cx)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double acc = Cx[e];
    bool hit = false;
    for (int64_t s = seg[e]; s < seg[e + 1]; s++)
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

struct SumBound : Bound
{
    int nops = 0, m = 0;
    CoalescePlan P;
    std::vector<int64_t> off;       // nops + 1 item offsets
    DBuf<int64_t> doff;
    DBuf<int> map;                  // item -> output entry
    DBuf<int> row;                  // item -> row (two-operand sums: the scaled fill)
    std::vector<int> dup;           // operand k repeats an entry
    std::vector<const double *> hxs;
    DBuf<const double *> xs;        // operand value pointers (fill-time upload target)
    DBuf<double> d;                 // 2 x m row scales (scaled fill)
};

// Coalescing plan of the operands' concatenated entries.
std::unique_ptr<SumBound> sum_plan(const char *fn, matrix *const *ops, int nops, bool keep_rows)
{
    if (nops < 1) api_exit(fn, "need at least one operand");
    const int m = ops[0]->m, n = ops[0]->n;
    auto b = std::make_unique<SumBound>();
    b->nops = nops;
    b->m = m;
    b->off.assign((size_t) nops + 1, 0);
    for (int k = 0; k < nops; k++)
    {
        if (ops[k]->m != m || ops[k]->n != n) api_exit(fn, "operand shapes differ");
        b->off[k + 1] = b->off[k] + ops[k]->nnz;
    }
    const int64_t items = b->off[nops];
    if (items > INT_MAX) api_exit(fn, "total operand nnz exceeds INT_MAX.");
    DBuf<int> orow(items), ocol(items);
    for (int k = 0; k < nops; k++)
    {
        SparseM *S = as_sparse(pd_csr_operand(ops[k]));
        sparse_ensure_dev(S);
        if (ops[k]->nnz == 0) continue;
        expand_rows(m, S->dp.p, orow.p + b->off[k], stream());
        d2d(ocol.p + b->off[k], S->di.p, (size_t) ops[k]->nnz * sizeof(int));
    }
    coalesce_plan(b->P, items, orow.p, ocol.p, m, n);
    b->doff = upload(b->off.data(), nops + 1);
    b->map.alloc(items);
    DBuf<int> ddup(nops);
    dzero(ddup.p, ddup.bytes());
    if (b->P.nnz > 0)
    {
        hipLaunchKernelGGL(k_sum_map, dim3(sum_grid(b->P.nnz)), dim3(256), 0, stream(), b->P.nnz,
                           b->P.seg.p, b->P.order.p, b->map.p);
        hipLaunchKernelGGL(k_sum_dups, dim3(sum_grid(b->P.nnz)), dim3(256), 0, stream(),
                           b->P.nnz, b->P.seg.p, b->P.order.p, b->doff.p, nops, ddup.p);
    }
    SDL_HIP(hipGetLastError());
    b->dup.assign(nops, 0);
    d2h(b->dup.data(), ddup.p, (size_t) nops * sizeof(int));
    if (keep_rows)
    {
        b->row = std::move(orow);
        b->d.alloc((size_t) 2 * m);
    }
    b->hxs.assign(nops, nullptr);
    b->xs.alloc(nops);
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
    const int nnz = checked_nnz(fn, b->P.nnz);
    CSR_matrix *c = new_CSR_matrix(m, n, nnz);
    write_structure(c, b->P);
    SparseM *C = as_sparse(make_sparse(c));
    sparse_ensure_dev(C);
    C->bound = std::move(b);
    return &C->base;
}

void check_operands(const char *fn, const SumBound *b, const matrix *const *ops, int nops)
{
    if (b->nops != nops) api_exit(fn, "operand count differs from the alloc");
    for (int k = 0; k < nops; k++)
        if (ops[k]->nnz != b->off[k + 1] - b->off[k])
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
    h2d(b->xs.p, b->hxs.data(), (size_t) nops * sizeof(const double *));
    if (b->P.nnz > 0)
    {
        if (d1)
        {
            if (!b->row.p) api_exit(fn, "output was not produced by sum_matrices_alloc");
            h2d(b->d.p, d1, (size_t) b->m * sizeof(double));
            h2d(b->d.p + b->m, d2, (size_t) b->m * sizeof(double));
            hipLaunchKernelGGL(k_sum_fill<true>, dim3(sum_grid(b->P.nnz)), dim3(256), 0, stream(),
                               b->P.nnz, b->P.seg.p, b->P.order.p, b->doff.p, nops, b->xs.p,
                               b->row.p, b->d.p, b->m, C->val.x.p);
        }
        else
        {
            hipLaunchKernelGGL(k_sum_fill<false>, dim3(sum_grid(b->P.nnz)), dim3(256), 0,
                               stream(), b->P.nnz, b->P.seg.p, b->P.order.p, b->doff.p, nops,
                               b->xs.p, (const int *) nullptr, (const double *) nullptr, b->m,
                               C->val.x.p);
        }
        SDL_HIP(hipGetLastError());
    }
    sparse_publish(C);
}

} // namespace

} // namespace sdl

using namespace sdl;

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
    const int nnz = checked_nnz(fn, b->P.nnz);
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
    write_structure(c, b->P);
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
    if (k < 0 || k >= b->nops) api_exit(fn, "operand index out of range");
    const int64_t lo = b->off[k], hi = b->off[k + 1];
    if (op->nnz != hi - lo) api_exit(fn, "operand nnz differs from the alloc");
    if (hi > lo)
    {
        // C's device values current (the caller may have written C->x)
        double *cx = const_cast<double *>(sparse_dev_values(C));
        const double *x =
            sparse_dev_values(as_sparse(pd_csr_operand_fresh(const_cast<matrix *>(op))));
        if (!b->dup[k])
            hipLaunchKernelGGL(k_sum_add, dim3(sum_grid(hi - lo)), dim3(256), 0, stream(),
                               hi - lo, b->map.p + lo, x, cx);
        else
            hipLaunchKernelGGL(k_sum_add_ordered, dim3(sum_grid(b->P.nnz)), dim3(256), 0,
                               stream(), b->P.nnz, b->P.seg.p, b->P.order.p, lo, hi, x, cx);
        SDL_HIP(hipGetLastError());
    }
    sparse_publish(C);
}

#pragma GCC visibility pop
} // extern "C"
