// Drop-in `matrix` handle layer, sparse kind and the kind-independent
// dispatchers (reference: src/matmul_dispatchers.c, src/matmul_kron_dispatchers.c,
// src/sparse/sparse_matrix.c, src/sparse/matrix_vstack.c).  Every alloc builds
// its output's structure on the GPU and binds what the fill needs; every fill
// refreshes the operands' device values (version-guarded), runs the GPU
// kernels, writes the output's host values and bumps its version once.
#include "api_matrix.h"
#include "mlp.h"
#include "api_pd.h"

#include <climits>
#include <cstring>

namespace sdl {

namespace {

__global__ void k_gather_ints(int64_t n, const int *__restrict__ idx, const int *__restrict__ in,
                              int *__restrict__ out)
{
    int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = in[idx[e]];
}

__global__ void k_offset_rowptr(int m, const int *__restrict__ Ap, int add, int *__restrict__ Cp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) Cp[i + 1] = Ap[i + 1] + add;
}

unsigned grid_for(int64_t n) { return (unsigned) std::max<int64_t>(1, (n + 255) / 256); }

void gather_ints(int64_t n, const int *idx, const int *in, int *out)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_ints, dim3(grid_for(n)), dim3(256), 0, stream(), n, idx, in, out);
    SDL_HIP(hipGetLastError());
}

// ---- bound states ---------------------------------------------------------
struct ProductBound : Bound
{
    ProductPlan P;
    DBuf<double> d;   // diagonal (fill-time upload target)
};

struct TransposeBound : Bound
{
    DBuf<int> perm;   // AT entry -> A position
};

struct GatherBound : Bound
{
    DBuf<int> src;    // C entry -> A position
};

struct ReduceBound : Bound
{
    CoalescePlan P;
};

struct ScaleBound : Bound
{
    DBuf<double> d;   // C = diag(d) A
};

struct VstackBound : Bound
{
    int n = 0;
    uint64_t C_seen = 0;
    std::vector<uint64_t> seen;
    std::vector<int64_t> nnz_off, row_off;
};

template <typename T> T *bound_as(SparseM *C, const char *fn)
{
    T *b = C->bound ? dynamic_cast<T *>(C->bound.get()) : nullptr;
    if (!b) api_exit(fn, "output was not produced by the matching alloc");
    return b;
}

// Build a host CSR from device int64 row pointers and column indices.
CSR_matrix *csr_from_device(const char *fn, int m, int n, const int64_t *dCp64, const int *dCi,
                            int64_t nnz64)
{
    int nnz = checked_nnz(fn, nnz64);
    CSR_matrix *C = new_CSR_matrix(m, n, nnz);
    DBuf<int> p32((size_t) m + 1);
    narrow_i64((int64_t) m + 1, dCp64, p32.p, stream());
    d2h(C->p, p32.p, ((size_t) m + 1) * sizeof(int));
    if (nnz > 0) d2h(C->i, dCi, (size_t) nnz * sizeof(int));
    return C;
}

// ---- vtable of the sparse kind -------------------------------------------
CSR_matrix *sparse_to_csr(matrix *self) { return as_sparse(self)->csr; }

matrix *sparse_row_gather_alloc(const matrix *self, const int *map, int m_out)
{
    SparseM *A = as_sparse(self);
    sparse_ensure_dev(A);
    DBuf<int> dmap = upload(map, m_out);
    DBuf<int64_t> cp((size_t) m_out + 1);
    int64_t nnz = row_gather_plan(m_out, dmap.p, A->dp.p, cp.p);
    if (nnz > INT_MAX)
        api_exit("sparse_row_gather_alloc", "gathered nnz exceeds INT_MAX.");
    auto b = std::make_unique<GatherBound>();
    b->src.alloc(nnz);
    row_gather_sources(m_out, dmap.p, A->dp.p, cp.p, nnz, b->src.p);
    DBuf<int> ci(nnz);
    gather_ints(nnz, b->src.p, A->di.p, ci.p);
    CSR_matrix *csr = csr_from_device("sparse_row_gather_alloc", m_out, self->n, cp.p, ci.p, nnz);
    SparseM *C = as_sparse(make_sparse(csr));
    sparse_ensure_dev(C);
    C->bound = std::move(b);
    return &C->base;
}

void sparse_row_gather_fill_values(const matrix *self, matrix *out)
{
    SparseM *A = as_sparse(self), *C = as_sparse(out);
    GatherBound *b = bound_as<GatherBound>(C, "sparse_row_gather_fill_values");
    gather_values(C->base.nnz, b->src.p, sparse_dev_values(A), C->val.x.p, stream());
    sparse_publish(C);
}

matrix *sparse_row_reduce_alloc(const matrix *self, const int *group, int m_out)
{
    SparseM *A = as_sparse(self);
    sparse_ensure_dev(A);
    sparse_ensure_rows(A);
    const int nnzA = self->nnz;
    DBuf<int> dgroup = upload(group, self->m);
    DBuf<int> orow(nnzA);
    gather_ints(nnzA, A->drow.p, dgroup.p, orow.p);
    auto b = std::make_unique<ReduceBound>();
    coalesce_plan(b->P, nnzA, orow.p, A->di.p, m_out, self->n);
    CSR_matrix *csr = csr_from_device("sparse_row_reduce_alloc", m_out, self->n, b->P.Cp.p,
                                      b->P.Ci.p, b->P.nnz);
    SparseM *C = as_sparse(make_sparse(csr));
    sparse_ensure_dev(C);
    C->bound = std::move(b);
    return &C->base;
}

void sparse_row_reduce_fill_values(const matrix *self, matrix *out)
{
    SparseM *A = as_sparse(self), *C = as_sparse(out);
    ReduceBound *b = bound_as<ReduceBound>(C, "sparse_row_reduce_fill_values");
    if (C->base.nnz > 0) coalesce_fill(b->P, sparse_dev_values(A), C->val.x.p, stream());
    sparse_publish(C);
}

// ---- sparse x sparse products -------------------------------------------
LView csc_view(SparseM *B)
{
    LView L;
    L.Lm = B->base.n;
    L.Lp = B->cp.p;
    L.Li = B->ci.p;
    L.Lperm = B->cperm.p;
    L.blocks = 1;
    return L;
}

LView rows_view(SparseM *B, int blocks)
{
    LView L;
    L.Lm = B->base.m;
    L.Lp = B->dp.p;
    L.Li = B->di.p;
    L.Lperm = nullptr;
    L.blocks = blocks;
    return L;
}

AView a_view(SparseM *A)
{
    AView v;
    v.Am = A->base.m;
    v.An = A->base.n;
    v.Ap = A->dp.p;
    v.Ai = A->di.p;
    return v;
}

matrix *product_alloc(const char *fn, const LView &L, SparseM *A, int d_len)
{
    auto b = std::make_unique<ProductBound>();
    product_plan(b->P, L, a_view(A), g_row_predictor ? SymMode::Predicted : SymMode::Exact,
                 g_row_predictor, true);
    CSR_matrix *csr = csr_from_device(fn, (int) b->P.m_out, A->base.n, b->P.Cp.p, b->P.Ci.p,
                                      b->P.nnz);
    b->d.alloc(std::max(d_len, 1));
    SparseM *C = as_sparse(make_sparse(csr));
    sparse_ensure_dev(C);
    C->bound = std::move(b);
    return &C->base;
}

void product_fill(const char *fn, const LView &L, const double *Lval, const double *d,
                  int d_len, SparseM *A, matrix *Cm)
{
    SparseM *C = as_sparse(Cm);
    ProductBound *b = bound_as<ProductBound>(C, fn);
    const double *Ax = sparse_dev_values(A);
    b->P.L = L;          // rebind to the operands given to this fill
    b->P.A = a_view(A);
    if (d) h2d(b->d.p, d, (size_t) d_len * sizeof(double));
    product_numeric(b->P, Lval, d ? b->d.p : nullptr, Ax, C->val.x.p, stream());
    sparse_publish(C);
}

} // namespace

// ---------------------------------------------------------------- helpers ----
SparseM *as_sparse(const matrix *A)
{
    if (!A || A->kind != MATRIX_SPARSE) api_exit("sparse kind", "expected a sparse matrix");
    return reinterpret_cast<SparseM *>(const_cast<matrix *>(A));
}

matrix *make_sparse(CSR_matrix *csr)
{
    SparseM *S = new SparseM();
    std::memset(&S->base, 0, sizeof(matrix));
    S->base.m = csr->m;
    S->base.n = csr->n;
    S->base.nnz = csr->nnz;
    S->base.x = csr->x;
    S->base.kind = MATRIX_SPARSE;
    S->base.to_csr = sparse_to_csr;
    S->base.row_gather_alloc = sparse_row_gather_alloc;
    S->base.row_gather_fill_values = sparse_row_gather_fill_values;
    S->base.row_reduce_alloc = sparse_row_reduce_alloc;
    S->base.row_reduce_fill_values = sparse_row_reduce_fill_values;
    S->csr = csr;
    return &S->base;
}

void sparse_ensure_dev(SparseM *S)
{
    if (S->dev) return;
    const CSR_matrix *A = S->csr;
    S->dp = upload(A->p, (int64_t) A->m + 1);
    S->di = upload(A->i, A->nnz);
    S->val.x.alloc(S->ext_values ? 0 : A->nnz);
    S->val.valid = false;
    S->dev = true;
}

void sparse_ensure_csc(SparseM *S)
{
    sparse_ensure_dev(S);
    if (S->has_csc) return;
    const CSR_matrix *A = S->csr;
    S->cp.alloc((size_t) A->n + 1);
    S->ci.alloc(A->nnz);
    S->cperm.alloc(A->nnz);
    transpose_structure(A->m, A->n, A->nnz, S->dp.p, S->di.p, S->cp.p, S->ci.p, S->cperm.p);
    S->has_csc = true;
}

void sparse_ensure_rows(SparseM *S)
{
    sparse_ensure_dev(S);
    if (S->drow.p) return;
    S->drow.alloc(S->csr->nnz);
    expand_rows(S->csr->m, S->dp.p, S->drow.p, stream());
    sync();
}

const double *sparse_dev_values(SparseM *S)
{
    if (S->ext_values) return S->ext_values(S);
    if (!S->dev) api_exit("fill", "operand has no device mirror (was it given to the alloc?)");
    if (!S->val.valid || S->val.seen != S->base.values_version)
    {
        h2d(S->val.x.p, S->csr->x, (size_t) S->csr->nnz * sizeof(double));
        S->val.seen = S->base.values_version;
        S->val.valid = true;
    }
    return S->val.x.p;
}

void sparse_publish(SparseM *S)
{
    d2h(S->csr->x, S->val.x.p, (size_t) S->csr->nnz * sizeof(double));
    matrix_values_changed(&S->base);
    S->val.seen = S->base.values_version;
    S->val.valid = true;
}

[[noreturn]] void unsupported(const char *fn, const matrix *A, const matrix *B)
{
    char msg[160];
    std::snprintf(msg, sizeof(msg), "operand kinds (%d, %d) not supported", A ? (int) A->kind : -1,
                  B ? (int) B->kind : -1);
    api_exit(fn, msg);
}

} // namespace sdl

using namespace sdl;

extern "C" {
#pragma GCC visibility push(default)

matrix *new_sparse_matrix(CSR_matrix *A) { return make_sparse(A); }

matrix *new_sparse_matrix_alloc(int m, int n, int nnz)
{
    return make_sparse(new_CSR_matrix(m, n, nnz));
}

void free_matrix(matrix *m)
{
    if (!m) return;
    if (m->kind == MATRIX_SPARSE)
    {
        SparseM *S = as_sparse(m);
        sync();
        free_CSR_matrix(S->csr);
        delete S;
        return;
    }
    pd_kind_free(m);
}

matrix *copy_sparsity_matrices_alloc(const matrix *A)
{
    if (A->kind != MATRIX_SPARSE) return pd_copy_sparsity(A);
    SparseM *S = as_sparse(A);
    sparse_ensure_rows(S);
    CSR_matrix *c = new_CSR_matrix(A->m, A->n, A->nnz);
    std::memcpy(c->p, S->csr->p, ((size_t) A->m + 1) * sizeof(int));
    std::memcpy(c->i, S->csr->i, (size_t) A->nnz * sizeof(int));
    SparseM *C = as_sparse(make_sparse(c));
    sparse_ensure_dev(C);
    auto b = std::make_unique<ScaleBound>();
    b->d.alloc(std::max(A->m, 1));
    C->bound = std::move(b);
    return &C->base;
}

void DA_matrices_fill_values(const double *d, const matrix *A, matrix *C)
{
    if (C->kind != A->kind) api_exit("DA_matrices_fill_values", "C has the wrong kind");
    if (A->kind != MATRIX_SPARSE)
    {
        pd_DA_fill(d, A, C);
        return;
    }
    SparseM *SA = as_sparse(A), *SC = as_sparse(C);
    ScaleBound *b = bound_as<ScaleBound>(SC, "DA_matrices_fill_values");
    sparse_ensure_rows(SA);  // no-op after copy_sparsity_matrices_alloc(A)
    h2d(b->d.p, d, (size_t) A->m * sizeof(double));
    scale_rows(A->nnz, SA->drow.p, sparse_dev_values(SA), b->d.p, SC->val.x.p, stream());
    sparse_publish(SC);
}

matrix *transpose_matrices_alloc(const matrix *A)
{
    if (A->kind != MATRIX_SPARSE) return pd_transpose_alloc(A);
    SparseM *S = as_sparse(A);
    sparse_ensure_dev(S);
    const CSR_matrix *a = S->csr;
    auto b = std::make_unique<TransposeBound>();
    DBuf<int> tp((size_t) a->n + 1), ti(a->nnz);
    b->perm.alloc(a->nnz);
    transpose_structure(a->m, a->n, a->nnz, S->dp.p, S->di.p, tp.p, ti.p, b->perm.p);
    CSR_matrix *t = new_CSR_matrix(a->n, a->m, a->nnz);
    d2h(t->p, tp.p, ((size_t) a->n + 1) * sizeof(int));
    d2h(t->i, ti.p, (size_t) a->nnz * sizeof(int));
    SparseM *C = as_sparse(make_sparse(t));
    sparse_ensure_dev(C);
    C->bound = std::move(b);
    return &C->base;
}

void transpose_matrices_fill_values(const matrix *A, matrix *AT)
{
    if (AT->kind != A->kind) api_exit("transpose_matrices_fill_values", "AT has the wrong kind");
    if (A->kind != MATRIX_SPARSE)
    {
        pd_transpose_fill(A, AT);
        return;
    }
    SparseM *S = as_sparse(A), *C = as_sparse(AT);
    TransposeBound *b = bound_as<TransposeBound>(C, "transpose_matrices_fill_values");
    gather_values(A->nnz, b->perm.p, sparse_dev_values(S), C->val.x.p, stream());
    sparse_publish(C);
}

// ------------------------------------------------------------ B^T A ----
matrix *BTA_matrices_alloc(matrix *A, matrix *B)
{
    if (A->kind == MATRIX_SPARSE && B->kind == MATRIX_SPARSE)
    {
        SparseM *SA = as_sparse(A), *SB = as_sparse(B);
        if (A->m != B->m) api_exit("BTA_matrices_alloc", "A and B row counts differ");
        sparse_ensure_dev(SA);
        sparse_ensure_csc(SB);
        return product_alloc("BTA_matrices_alloc", csc_view(SB), SA, A->m);
    }
    return pd_BTA_alloc(A, B);
}

static void btda_fill(const char *fn, matrix *A, const double *d, matrix *B, matrix *C)
{
    if (A->kind == MATRIX_SPARSE && B->kind == MATRIX_SPARSE)
    {
        if (C->kind != MATRIX_SPARSE) api_exit(fn, "C has the wrong kind");
        SparseM *SA = as_sparse(A), *SB = as_sparse(B);
        sparse_ensure_csc(SB);
        product_fill(fn, csc_view(SB), sparse_dev_values(SB), d, A->m, SA, C);
        return;
    }
    pd_BTDA_fill(A, d, B, C);
}

void BTDA_matrices_fill_values(matrix *A, const double *d, matrix *B, matrix *C)
{
    btda_fill("BTDA_matrices_fill_values", A, d, B, C);
}

void BTA_matrices_fill_values(matrix *A, matrix *B, matrix *C)
{
    btda_fill("BTA_matrices_fill_values", A, nullptr, B, C);
}

// ------------------------------------------------------------ A^T D A ----
matrix *ATA_matrices_alloc(matrix *A)
{
    if (A->kind != MATRIX_SPARSE) return pd_ATA_alloc(A);
    SparseM *S = as_sparse(A);
    sparse_ensure_csc(S);
    return product_alloc("ATA_matrices_alloc", csc_view(S), S, A->m);
}

void ATDA_matrices_fill_values(const matrix *A, const double *d, matrix *C)
{
    if (C->kind != A->kind) api_exit("ATDA_matrices_fill_values", "C has the wrong kind");
    if (A->kind != MATRIX_SPARSE)
    {
        pd_ATDA_fill(A, d, C);
        return;
    }
    SparseM *S = as_sparse(A);
    sparse_ensure_csc(S);
    product_fill("ATDA_matrices_fill_values", csc_view(S), sparse_dev_values(S), d, A->m, S, C);
}

// --------------------------------------------------------------- B A ----
matrix *BA_matrices_alloc(const matrix *B, matrix *A)
{
    if (B->kind == MATRIX_STACKED_PD) api_exit("BA_matrices_alloc", "a stacked_pd B is not supported");
    if (B->kind == MATRIX_SPARSE && A->kind == MATRIX_SPARSE)
    {
        SparseM *SB = as_sparse(B), *SA = as_sparse(A);
        if (B->n != A->m) api_exit("BA_matrices_alloc", "inner dimensions differ");
        sparse_ensure_dev(SB);
        sparse_ensure_dev(SA);
        return product_alloc("BA_matrices_alloc", rows_view(SB, 1), SA, 0);
    }
    return pd_BA_alloc(B, A);
}

void BA_matrices_fill_values(const matrix *B, const matrix *A, matrix *C)
{
    if (B->kind == MATRIX_STACKED_PD) api_exit("BA_matrices_fill_values", "a stacked_pd B is not supported");
    if (B->kind == MATRIX_SPARSE && A->kind == MATRIX_SPARSE)
    {
        if (C->kind != MATRIX_SPARSE) api_exit("BA_matrices_fill_values", "C has the wrong kind");
        SparseM *SB = as_sparse(B), *SA = as_sparse(A);
        product_fill("BA_matrices_fill_values", rows_view(SB, 1), sparse_dev_values(SB), nullptr,
                     0, SA, C);
        return;
    }
    pd_BA_fill(B, A, C);
}

// -------------------------------------------------------------- kron ----
void BA_kron_vec_matrices(const matrix *A, const double *x, double *y, int p)
{
    if (A->kind == MATRIX_STACKED_PD) api_exit("BA_kron_vec_matrices", "stacked_pd A not supported");
    if (A->kind == MATRIX_PERMUTED_DENSE)
    {
        pd_kron_vec(A, x, y, p);
        return;
    }
    SparseM *S = as_sparse(A);
    sparse_ensure_dev(S);
    DBuf<double> dx = upload(x, (int64_t) A->n * p);
    DBuf<double> dy((size_t) A->m * p);
    kron_spmv(A->m, A->n, S->dp.p, S->di.p, sparse_dev_values(S), dx.p, dy.p, p, stream());
    d2h(y, dy.p, (size_t) A->m * p * sizeof(double));
}

matrix *BA_kron_matrices_alloc(const matrix *A, int p, matrix *J)
{
    if (p == 1) return BA_matrices_alloc(A, J);
    if (A->kind == MATRIX_PERMUTED_DENSE) return pd_dense_kron_alloc(A, p, J, false);
    if (A->kind != MATRIX_SPARSE) api_exit("BA_kron_matrices_alloc", "stacked_pd A not supported");
    if (J->m != p * A->n) api_exit("BA_kron_matrices_alloc", "J->m must equal p * A->n");
    SparseM *SA = as_sparse(A);
    sparse_ensure_dev(SA);
    SparseM *SJ = as_sparse(pd_csr_operand(J));
    sparse_ensure_dev(SJ);
    return product_alloc("BA_kron_matrices_alloc", rows_view(SA, p), SJ, 0);
}

void BA_kron_matrices_fill_values(const matrix *A, int p, matrix *J, matrix *C)
{
    if (p == 1)
    {
        BA_matrices_fill_values(A, J, C);
        return;
    }
    if (A->kind == MATRIX_PERMUTED_DENSE)
    {
        pd_dense_kron_fill(A, p, J, C, false);
        return;
    }
    SparseM *SA = as_sparse(A);
    SparseM *SJ = as_sparse(pd_csr_operand(J));
    product_fill("BA_kron_matrices_fill_values", rows_view(SA, p), sparse_dev_values(SA), nullptr,
                 0, SJ, C);
}

matrix *BA_dense_kron_I_matrices_alloc(const matrix *A, int m, matrix *J)
{
    return pd_dense_kron_alloc(A, m, J, true);
}

void BA_dense_kron_I_matrices_fill_values(const matrix *A, int m, const matrix *J, matrix *C)
{
    pd_dense_kron_fill(A, m, const_cast<matrix *>(J), C, true);
}

// ------------------------------------------------------------- vstack ----
matrix *vstack_matrices_alloc(matrix *const *ops, int n)
{
    if (n < 1) api_exit("vstack_matrices_alloc", "need at least one operand");
    const int ncols = ops[0]->n;
    int64_t m_total = 0, nnz_total = 0;
    for (int k = 0; k < n; k++)
    {
        if (ops[k]->n != ncols) api_exit("vstack_matrices_alloc", "column counts differ");
        m_total += ops[k]->m;
        nnz_total += ops[k]->nnz;
    }
    if (m_total > INT_MAX || nnz_total > INT_MAX)
        api_exit("vstack_matrices_alloc", "stacked size exceeds INT_MAX.");
    auto b = std::make_unique<VstackBound>();
    b->n = n;
    b->seen.assign(n, 0);
    DBuf<int> cp((size_t) m_total + 1), ci(nnz_total);
    dzero(cp.p, sizeof(int));
    int64_t row_off = 0, nnz_off = 0;
    for (int k = 0; k < n; k++)
    {
        SparseM *S = as_sparse(pd_csr_operand(ops[k]));
        sparse_ensure_dev(S);
        if (ops[k]->m > 0)
            hipLaunchKernelGGL(k_offset_rowptr, dim3(grid_for(ops[k]->m)), dim3(256), 0, stream(),
                               ops[k]->m, S->dp.p, (int) nnz_off, cp.p + row_off);
        d2d(ci.p + nnz_off, S->di.p, (size_t) ops[k]->nnz * sizeof(int));
        b->row_off.push_back(row_off);
        b->nnz_off.push_back(nnz_off);
        row_off += ops[k]->m;
        nnz_off += ops[k]->nnz;
    }
    SDL_HIP(hipGetLastError());
    CSR_matrix *c = new_CSR_matrix((int) m_total, ncols, (int) nnz_total);
    d2h(c->p, cp.p, ((size_t) m_total + 1) * sizeof(int));
    d2h(c->i, ci.p, (size_t) nnz_total * sizeof(int));
    SparseM *C = as_sparse(make_sparse(c));
    sparse_ensure_dev(C);
    b->C_seen = C->base.values_version - 1;  // stale: the first fill copies everything
    C->bound = std::move(b);
    return &C->base;
}

void vstack_matrices_fill_values(matrix *const *ops, int n, matrix *Cm)
{
    SparseM *C = as_sparse(Cm);
    VstackBound *b = bound_as<VstackBound>(C, "vstack_matrices_fill_values");
    if (b->n != n) api_exit("vstack_matrices_fill_values", "operand count differs from the alloc");
    const bool copy_all = b->C_seen != Cm->values_version;
    bool changed = false;
    for (int k = 0; k < n; k++)
    {
        matrix *op = ops[k];
        if ((copy_all || b->seen[k] != op->values_version) && op->nnz > 0)
        {
            SparseM *S = as_sparse(pd_csr_operand_fresh(op));
            d2d(C->val.x.p + b->nnz_off[k], sparse_dev_values(S),
                (size_t) op->nnz * sizeof(double));
            changed = true;
        }
        b->seen[k] = op->values_version;
    }
    if (changed)
    {
        sparse_publish(C);
    }
    b->C_seen = Cm->values_version;
}

#pragma GCC visibility pop
} // extern "C"
