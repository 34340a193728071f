// permuted_dense / stacked_pd kinds: constructors, device mirrors, CSR views,
// signature coalescing, and the structural operations (copy_sparsity, DA,
// transpose, row_gather, row_reduce).  Reference: src/permuted_dense/
// permuted_dense.c, src/stacked_pd/stacked_pd.c, stacked_pd_coalesce.c,
// stacked_pd_linalg.c:30-193.
#include "pd_internal.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <map>

namespace sdl {

// -------------------------------------------------------------- helpers ----
bool has_overlap(const std::vector<int> &a, const std::vector<int> &b)
{
    size_t i = 0, j = 0;
    while (i < a.size() && j < b.size())
    {
        if (a[i] == b[j]) return true;
        if (a[i] < b[j]) i++;
        else j++;
    }
    return false;
}

int sorted_intersect(const std::vector<int> &a, const std::vector<int> &b, std::vector<int> &ia,
                     std::vector<int> &ib)
{
    ia.clear();
    ib.clear();
    size_t i = 0, j = 0;
    while (i < a.size() && j < b.size())
    {
        if (a[i] == b[j])
        {
            ia.push_back((int) i++);
            ib.push_back((int) j++);
        }
        else if (a[i] < b[j]) i++;
        else j++;
    }
    return (int) ia.size();
}

std::vector<int> sorted_union(const std::vector<const std::vector<int> *> &arrs)
{
    std::vector<int> out;
    for (auto *a : arrs) out.insert(out.end(), a->begin(), a->end());
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

PDM *as_pd(const matrix *A)
{
    if (!A || A->kind != MATRIX_PERMUTED_DENSE) api_exit("permuted_dense", "expected a permuted_dense");
    return reinterpret_cast<PDM *>(const_cast<matrix *>(A));
}

SPDM *as_spd(const matrix *A)
{
    if (!A || A->kind != MATRIX_STACKED_PD) api_exit("stacked_pd", "expected a stacked_pd");
    return reinterpret_cast<SPDM *>(const_cast<matrix *>(A));
}

// ------------------------------------------------------- device values ----
const double *spd_vals(SPDM *S)
{
    if (!S->dvalid || S->dseen != S->base.values_version)
    {
        h2d(S->dx.p, S->x, (size_t) S->base.nnz * sizeof(double));
        S->dseen = S->base.values_version;
        S->dvalid = true;
    }
    return S->dx.p;
}

const double *pd_vals(PDM *A)
{
    if (A->parent) return spd_vals(A->parent) + A->poff;
    if (!A->dvalid || A->dseen != A->base.values_version)
    {
        h2d(A->dX.p, A->X, (size_t) A->base.nnz * sizeof(double));
        A->dseen = A->base.values_version;
        A->dvalid = true;
    }
    return A->dX.p;
}

const double *mat_vals(matrix *A)
{
    if (A->kind == MATRIX_PERMUTED_DENSE) return pd_vals(as_pd(A));
    if (A->kind == MATRIX_STACKED_PD) return spd_vals(as_spd(A));
    return sparse_dev_values(as_sparse(A));
}

double *out_buf(matrix *C)
{
    if (C->kind == MATRIX_PERMUTED_DENSE)
    {
        PDM *P = as_pd(C);
        if (P->parent) return P->parent->dx.p + P->poff;
        return P->dX.p;
    }
    if (C->kind == MATRIX_STACKED_PD) return as_spd(C)->dx.p;
    return as_sparse(C)->val.x.p;
}

void out_publish(matrix *C)
{
    if (C->kind == MATRIX_PERMUTED_DENSE)
    {
        PDM *P = as_pd(C);
        if (P->parent) api_exit("fill", "output is a block of a stack");
        d2h(P->X, P->dX.p, (size_t) C->nnz * sizeof(double));
        matrix_values_changed(C);
        P->dseen = C->values_version;
        P->dvalid = true;
        return;
    }
    if (C->kind == MATRIX_STACKED_PD)
    {
        SPDM *S = as_spd(C);
        d2h(S->x, S->dx.p, (size_t) C->nnz * sizeof(double));
        matrix_values_changed(C);
        S->dseen = C->values_version;
        S->dvalid = true;
        return;
    }
    sparse_publish(as_sparse(C));
}

// ------------------------------------------------------------ CSR views ----
namespace {

const double *pd_view_vals(SparseM *V) { return pd_vals(reinterpret_cast<PDM *>(V->ext_owner)); }

const double *spd_view_vals(SparseM *V)
{
    SPDM *S = reinterpret_cast<SPDM *>(V->ext_owner);
    if (!S->csr_valid || S->csr_seen != S->base.values_version)
    {
        gather_values(S->base.nnz, S->csr_gather.idx.p, spd_vals(S), V->val.x.p, stream());
        S->csr_seen = S->base.values_version;
        S->csr_valid = true;
    }
    return V->val.x.p;
}

// pattern of a union of row-disjoint blocks (pd_blocks_csr_pattern,
// permuted_dense.c:344-367) and, per CSR position, the block-major position
void blocks_csr_pattern(const std::vector<PDM *> &blocks, const std::vector<int64_t> &boff, int m,
                        int *Cp, int *Ci, std::vector<int> *pos)
{
    std::memset(Cp, 0, ((size_t) m + 1) * sizeof(int));
    for (PDM *b : blocks)
        for (int ii = 0; ii < b->m0; ii++) Cp[b->row_perm[ii] + 1] = b->n0;
    for (int r = 0; r < m; r++) Cp[r + 1] += Cp[r];
    if (pos) pos->assign(Cp[m], 0);
    for (size_t k = 0; k < blocks.size(); k++)
    {
        PDM *b = blocks[k];
        for (int ii = 0; ii < b->m0; ii++)
        {
            int start = Cp[b->row_perm[ii]];
            for (int jj = 0; jj < b->n0; jj++)
            {
                Ci[start + jj] = b->col_perm[jj];
                if (pos) (*pos)[start + jj] = (int) (boff[k] + (int64_t) ii * b->n0 + jj);
            }
        }
    }
}

matrix *pd_view(PDM *A)
{
    if (A->csrv) return A->csrv;
    CSR_matrix *c = new_CSR_matrix(A->base.m, A->base.n, A->base.nnz);
    std::vector<PDM *> one{A};
    std::vector<int64_t> off{0};
    blocks_csr_pattern(one, off, A->base.m, c->p, c->i, nullptr);
    host_free(c->x);
    c->x = A->X;  // values alias X (row-major block = CSR order)
    SparseM *S = as_sparse(make_sparse(c));
    S->ext_values = pd_view_vals;
    S->ext_owner = A;
    sparse_ensure_dev(S);
    A->csrv = &S->base;
    return A->csrv;
}

matrix *spd_view(SPDM *A)
{
    if (A->csrv) return A->csrv;
    CSR_matrix *c = new_CSR_matrix(A->base.m, A->base.n, A->base.nnz);
    std::vector<int> pos;
    blocks_csr_pattern(A->blocks, A->boff, A->base.m, c->p, c->i, &pos);
    SparseM *S = as_sparse(make_sparse(c));
    S->ext_values = spd_view_vals;
    S->ext_owner = A;
    sparse_ensure_dev(S);
    S->val.x.alloc(A->base.nnz);
    A->csr_gather.n = A->base.nnz;
    A->csr_gather.idx = upload(pos.data(), (int64_t) pos.size());
    A->csr_valid = false;
    A->csrv = &S->base;
    return A->csrv;
}

void free_view(matrix *v, bool alias_x)
{
    if (!v) return;
    SparseM *S = as_sparse(v);
    if (alias_x) S->csr->x = nullptr;
    sync();
    free_CSR_matrix(S->csr);
    delete S;
}

CSR_matrix *pd_to_csr(matrix *self)
{
    PDM *A = as_pd(self);
    return as_sparse(pd_view(A))->csr;
}

CSR_matrix *spd_to_csr(matrix *self)
{
    SPDM *A = as_spd(self);
    SparseM *V = as_sparse(spd_view(A));
    if (!A->csr_valid || A->csr_seen != A->base.values_version)
    {
        spd_view_vals(V);
        d2h(V->csr->x, V->val.x.p, (size_t) A->base.nnz * sizeof(double));
    }
    return V->csr;
}

// ------------------------------------------------------------ bound states ----
struct GatherB : Bound
{
    DBuf<int> perm;  // output position -> source position
};

struct PullB : Bound
{
    PullList pull;
};

struct TwoStageB : Bound
{
    PullList first;    // source -> raw partials
    DBuf<double> raw;
    PullList second;   // raw -> output
};

template <typename T> T *bound_of(matrix *C, const char *fn)
{
    std::unique_ptr<Bound> *b = C->kind == MATRIX_PERMUTED_DENSE ? &as_pd(C)->bound
                                : C->kind == MATRIX_STACKED_PD  ? &as_spd(C)->bound
                                                                : &as_sparse(C)->bound;
    T *t = *b ? dynamic_cast<T *>(b->get()) : nullptr;
    if (!t) api_exit(fn, "output was not produced by the matching alloc");
    return t;
}

void set_bound(matrix *C, std::unique_ptr<Bound> b)
{
    if (C->kind == MATRIX_PERMUTED_DENSE) as_pd(C)->bound = std::move(b);
    else if (C->kind == MATRIX_STACKED_PD) as_spd(C)->bound = std::move(b);
    else as_sparse(C)->bound = std::move(b);
}

Blk blk_of(const PDM *A)
{
    Blk b;
    b.m = A->base.m;
    b.n = A->base.n;
    b.rp = A->row_perm;
    b.cp = A->col_perm;
    return b;
}

// Build an output of `kind` from block structures: a PD (single block) or a
// stack.  Values are left for the fill.
matrix *make_output_blocks(bool as_stack, int m, int n, const std::vector<Blk> &blocks,
                           const std::vector<int> &src_p, const std::vector<int> &src_idx)
{
    if (!as_stack)
    {
        if (blocks.size() != 1) api_exit("internal", "expected one block");
        return &new_pdm(blocks[0], nullptr)->base;
    }
    std::vector<PDM *> pds;
    for (const Blk &b : blocks) pds.push_back(new_pdm(b, nullptr));
    return &new_spdm(m, n, pds, src_p, src_idx)->base;
}

// block-major offsets of a set of block structures
std::vector<long long> offsets_of(const std::vector<Blk> &blocks, long long *total)
{
    std::vector<long long> off;
    long long t = 0;
    for (const Blk &b : blocks)
    {
        off.push_back(t);
        t += b.size();
    }
    if (total) *total = t;
    return off;
}

// ------------------------------------------------------------- vtable ----
matrix *pd_row_gather_alloc(const matrix *self, const int *map, int m_out);
void pd_row_gather_fill(const matrix *self, matrix *out);
matrix *pd_row_reduce_alloc(const matrix *self, const int *group, int m_out);
void pd_row_reduce_fill(const matrix *self, matrix *out);

void wire_pd(PDM *P)
{
    P->base.kind = MATRIX_PERMUTED_DENSE;
    P->base.to_csr = pd_to_csr;
    P->base.row_gather_alloc = pd_row_gather_alloc;
    P->base.row_gather_fill_values = pd_row_gather_fill;
    P->base.row_reduce_alloc = pd_row_reduce_alloc;
    P->base.row_reduce_fill_values = pd_row_reduce_fill;
}

void wire_spd(SPDM *S)
{
    S->base.kind = MATRIX_STACKED_PD;
    S->base.to_csr = spd_to_csr;
    S->base.row_gather_alloc = pd_row_gather_alloc;
    S->base.row_gather_fill_values = pd_row_gather_fill;
    S->base.row_reduce_alloc = pd_row_reduce_alloc;
    S->base.row_reduce_fill_values = pd_row_reduce_fill;
}

// the blocks of a PD (itself) or an SPD, with offsets into its value buffer
void blocks_of(const matrix *A, std::vector<PDM *> &bl, std::vector<long long> &off)
{
    bl.clear();
    off.clear();
    if (A->kind == MATRIX_PERMUTED_DENSE)
    {
        bl.push_back(as_pd(A));
        off.push_back(0);
    }
    else
    {
        SPDM *S = as_spd(A);
        bl = S->blocks;
        for (int64_t o : S->boff) off.push_back(o);
    }
}

// row_gather of PD (permuted_dense.c:72-112) and SPD (stacked_pd.c:161-194):
// output rows = map positions whose source row is dense; one copy per value.
matrix *pd_row_gather_alloc(const matrix *self, const int *map, int m_out)
{
    std::vector<PDM *> bl;
    std::vector<long long> off;
    blocks_of(self, bl, off);
    std::vector<Blk> out;
    std::vector<int> src_idx;
    std::vector<int> perm;
    for (size_t k = 0; k < bl.size(); k++)
    {
        PDM *B = bl[k];
        Blk b;
        b.m = m_out;
        b.n = self->n;
        b.cp = B->col_perm;
        std::vector<int> src_rows;
        for (int i = 0; i < m_out; i++)
        {
            int ii = B->row_inv[map[i]];
            if (ii >= 0)
            {
                b.rp.push_back(i);
                src_rows.push_back(ii);
            }
        }
        if (self->kind == MATRIX_STACKED_PD && b.size() == 0) continue;  // spd_map_filter_blocks
        for (size_t r = 0; r < src_rows.size(); r++)
            for (int jj = 0; jj < B->n0; jj++)
                perm.push_back((int) (off[k] + (long long) src_rows[r] * B->n0 + jj));
        out.push_back(b);
        src_idx.push_back((int) k);
    }
    std::vector<int> src_p(out.size() + 1);
    for (size_t k = 0; k <= out.size(); k++) src_p[k] = (int) k;
    matrix *C = make_output_blocks(self->kind == MATRIX_STACKED_PD, m_out, self->n, out, src_p,
                                   src_idx);
    auto b = std::make_unique<GatherB>();
    b->perm = upload(perm.data(), (int64_t) perm.size());
    set_bound(C, std::move(b));
    return C;
}

void pd_row_gather_fill(const matrix *self, matrix *out)
{
    GatherB *b = bound_of<GatherB>(out, "row_gather_fill_values");
    gather_values(out->nnz, b->perm.p, mat_vals(const_cast<matrix *>(self)), out_buf(out), stream());
    out_publish(out);
}

// row_reduce of PD (permuted_dense.c:152-209): output rows = groups hit by the
// dense rows, values summed in ascending source row from 0.  SPD
// (stacked_pd_linalg.c:176-193): per-block partials, then the coalesce-
// accumulate of the partials in block order.
matrix *pd_row_reduce_alloc(const matrix *self, const int *group, int m_out)
{
    std::vector<PDM *> bl;
    std::vector<long long> off;
    blocks_of(self, bl, off);
    std::vector<Blk> raw;
    std::vector<int> seg1{0}, idx1;
    for (size_t k = 0; k < bl.size(); k++)
    {
        PDM *B = bl[k];
        std::vector<int> g2o(m_out, -1);
        for (int ii = 0; ii < B->m0; ii++) g2o[group[B->row_perm[ii]]] = 0;
        Blk b;
        b.m = m_out;
        b.n = self->n;
        b.cp = B->col_perm;
        for (int g = 0; g < m_out; g++)
            if (g2o[g] >= 0)
            {
                g2o[g] = (int) b.rp.size();
                b.rp.push_back(g);
            }
        // per partial cell (g', jj): the source dense rows with that group, ascending
        std::vector<std::vector<int>> rows_of(b.rp.size());
        for (int ii = 0; ii < B->m0; ii++) rows_of[g2o[group[B->row_perm[ii]]]].push_back(ii);
        for (size_t r = 0; r < b.rp.size(); r++)
            for (int jj = 0; jj < B->n0; jj++)
            {
                for (int ii : rows_of[r]) idx1.push_back((int) (off[k] + (long long) ii * B->n0 + jj));
                seg1.push_back((int) idx1.size());
            }
        raw.push_back(b);
    }
    if (self->kind == MATRIX_PERMUTED_DENSE)
    {
        std::vector<int> src_p{0, 1}, src_idx{0};
        matrix *C = make_output_blocks(false, m_out, self->n, raw, src_p, src_idx);
        auto b = std::make_unique<PullB>();
        pull_build(b->pull, seg1, idx1);
        set_bound(C, std::move(b));
        return C;
    }
    long long raw_total = 0;
    std::vector<long long> raw_off = offsets_of(raw, &raw_total);
    Coalesced co = coalesce_blocks(m_out, self->n, raw, raw_off);
    matrix *C = make_output_blocks(true, m_out, self->n, co.out, co.src_p, co.src_idx);
    auto b = std::make_unique<TwoStageB>();
    pull_build(b->first, seg1, idx1);
    b->raw.alloc(raw_total);
    pull_build(b->second, co.seg, co.idx);
    set_bound(C, std::move(b));
    return C;
}

void pd_row_reduce_fill(const matrix *self, matrix *out)
{
    const double *src = mat_vals(const_cast<matrix *>(self));
    if (out->kind == MATRIX_PERMUTED_DENSE)
    {
        PullB *b = bound_of<PullB>(out, "row_reduce_fill_values");
        pull_sum(b->pull, src, out_buf(out), stream());
    }
    else
    {
        TwoStageB *b = bound_of<TwoStageB>(out, "row_reduce_fill_values");
        pull_sum(b->first, src, b->raw.p, stream());
        pull_sum(b->second, b->raw.p, out_buf(out), stream());
    }
    out_publish(out);
}

} // namespace

// --------------------------------------------------------- constructors ----
PDM *new_pdm(const Blk &b, const double *X)
{
    const int m0 = b.m0(), n0 = b.n0();
    for (int ii = 1; ii < m0; ii++)
        if (b.rp[ii] <= b.rp[ii - 1]) api_exit("new_permuted_dense", "row_perm must be strictly increasing");
    for (int jj = 1; jj < n0; jj++)
        if (b.cp[jj] <= b.cp[jj - 1]) api_exit("new_permuted_dense", "col_perm must be strictly increasing");
    if (m0 > 0 && (b.rp[0] < 0 || b.rp[m0 - 1] >= b.m)) api_exit("new_permuted_dense", "row_perm out of range");
    if (n0 > 0 && (b.cp[0] < 0 || b.cp[n0 - 1] >= b.n)) api_exit("new_permuted_dense", "col_perm out of range");
    if (m0 > 0 && n0 > INT_MAX / m0)
        api_exit("new_permuted_dense", "dense block m0 * n0 exceeds INT_MAX.");
    PDM *P = new PDM();
    std::memset(&P->base, 0, sizeof(matrix));
    P->base.m = b.m;
    P->base.n = b.n;
    P->base.nnz = m0 * n0;
    wire_pd(P);
    P->m0 = m0;
    P->n0 = n0;
    P->row_perm = b.rp;
    P->col_perm = b.cp;
    P->row_inv.assign(b.m, -1);
    P->col_inv.assign(b.n, -1);
    for (int ii = 0; ii < m0; ii++) P->row_inv[b.rp[ii]] = ii;
    for (int jj = 0; jj < n0; jj++) P->col_inv[b.cp[jj]] = jj;
    P->X = host_doubles((int64_t) m0 * n0);
    if (X && m0 * n0 > 0) std::memcpy(P->X, X, (size_t) m0 * n0 * sizeof(double));
    P->base.x = P->X;
    P->dX.alloc((size_t) m0 * n0);
    return P;
}

SPDM *new_spdm(int m, int n, const std::vector<PDM *> &blocks, const std::vector<int> &src_p,
               const std::vector<int> &src_idx)
{
    SPDM *S = new SPDM();
    std::memset(&S->base, 0, sizeof(matrix));
    S->base.m = m;
    S->base.n = n;
    long long nnz = 0;
    for (PDM *b : blocks) nnz += b->base.nnz;
    if (nnz > INT_MAX) api_exit("new_stacked_pd", "nnz exceeds INT_MAX.");
    S->base.nnz = (int) nnz;
    wire_spd(S);
    S->blocks = blocks;
    S->src_p = src_p;
    S->src_idx = src_idx;
    S->x = host_doubles(nnz);
    S->base.x = S->x;
    S->dx.alloc(nnz);
    long long off = 0;
    for (PDM *b : blocks)
    {
        S->boff.push_back(off);
        int bn = b->base.nnz;
        if (bn > 0) std::memcpy(S->x + off, b->X, (size_t) bn * sizeof(double));
        if (!b->parent) host_free(b->X);
        b->X = S->x + off;
        b->base.x = b->X;
        b->parent = S;
        b->poff = off;
        b->dX.reset();
        if (b->csrv) as_sparse(b->csrv)->csr->x = b->X;
        off += bn;
    }
    return S;
}

void pd_kind_free(matrix *m)
{
    if (m->kind == MATRIX_PERMUTED_DENSE)
    {
        PDM *P = as_pd(m);
        sync();
        free_view(P->csrv, true);
        if (!P->parent) host_free(P->X);
        delete P;
        return;
    }
    SPDM *S = as_spd(m);
    sync();
    for (PDM *b : S->blocks)
    {
        free_view(b->csrv, true);
        delete b;
    }
    free_view(S->csrv, false);
    host_free(S->x);
    delete S;
}

matrix *pd_csr_operand(matrix *M)
{
    if (M->kind == MATRIX_SPARSE) return M;
    if (M->kind == MATRIX_PERMUTED_DENSE) return pd_view(as_pd(M));
    return spd_view(as_spd(M));
}

matrix *pd_csr_operand_fresh(matrix *M) { return pd_csr_operand(M); }

// ------------------------------------------------- signature coalescing ----
Coalesced coalesce_blocks(int m, int n, const std::vector<Blk> &raw,
                          const std::vector<long long> &raw_off, bool lists)
{
    Coalesced co;
    // blocks touching each row, ascending block id (stacked_pd_coalesce.c:101-125)
    std::vector<std::vector<int>> row_blocks(m);
    for (size_t k = 0; k < raw.size(); k++)
        for (int r : raw[k].rp) row_blocks[r].push_back((int) k);
    // signatures in row-scan order (:146-166)
    std::map<std::vector<int>, int> sig_id;
    std::vector<std::vector<int>> sigs;
    std::vector<std::vector<int>> sig_rows;
    for (int r = 0; r < m; r++)
    {
        if (row_blocks[r].empty()) continue;
        auto it = sig_id.find(row_blocks[r]);
        int s;
        if (it == sig_id.end())
        {
            s = (int) sigs.size();
            sig_id.emplace(row_blocks[r], s);
            sigs.push_back(row_blocks[r]);
            sig_rows.emplace_back();
        }
        else s = it->second;
        sig_rows[s].push_back(r);
    }
    co.src_p.push_back(0);
    for (size_t s = 0; s < sigs.size(); s++)
    {
        Blk b;
        b.m = m;
        b.n = n;
        b.rp = sig_rows[s];
        std::vector<const std::vector<int> *> arrs;
        for (int k : sigs[s]) arrs.push_back(&raw[k].cp);
        b.cp = sorted_union(arrs);
        co.out.push_back(b);
        for (int k : sigs[s]) co.src_idx.push_back(k);
        co.src_p.push_back((int) co.src_idx.size());
    }
    // per output cell, the raw cells feeding it in raw order (:326-371)
    if (!lists) return co;
    co.seg.push_back(0);
    for (size_t s = 0; s < sigs.size(); s++)
    {
        const Blk &o = co.out[s];
        std::vector<std::vector<int>> rinv, cinv;
        for (int k : sigs[s])
        {
            std::vector<int> ri(m, -1), ci(n, -1);
            for (int ii = 0; ii < raw[k].m0(); ii++) ri[raw[k].rp[ii]] = ii;
            for (int jj = 0; jj < raw[k].n0(); jj++) ci[raw[k].cp[jj]] = jj;
            rinv.push_back(std::move(ri));
            cinv.push_back(std::move(ci));
        }
        for (int r : o.rp)
            for (int c : o.cp)
            {
                for (size_t t = 0; t < sigs[s].size(); t++)
                {
                    int k = sigs[s][t];
                    int ii = rinv[t][r], jj = cinv[t][c];
                    if (ii >= 0 && jj >= 0)
                        co.idx.push_back((int) (raw_off[k] + (long long) ii * raw[k].n0() + jj));
                }
                co.seg.push_back((int) co.idx.size());
            }
    }
    return co;
}

bool rowc_eligible(int m, int n, const std::vector<Blk> &raw)
{
    if (g_coalesce_mode != 0 || n <= 0 || n > ROWC_MAX_N) return false;
    std::vector<int> seen_r(std::max(m, 1), -1), seen_c(n, -1);
    for (size_t k = 0; k < raw.size(); k++)
    {
        for (int r : raw[k].rp)
        {
            if (r < 0 || r >= m || seen_r[r] == (int) k) return false;
            seen_r[r] = (int) k;
        }
        for (int c : raw[k].cp)
        {
            if (c < 0 || c >= n || seen_c[c] == (int) k) return false;
            seen_c[c] = (int) k;
        }
    }
    return true;
}

// Output rows in output order (signature blocks in catalogue order, their
// rows in row-scan order); each row's sources are its signature's raw blocks
// in ascending order, source = that block's row holding r.
void rowc_build(RowCoalesce &P, int m, int n, const std::vector<Blk> &raw,
                const std::vector<long long> &raw_off, const Coalesced &co)
{
    // local row of every (raw block, global row) pair, by global row
    std::vector<std::vector<std::pair<int, int>>> at(m);
    for (size_t k = 0; k < raw.size(); k++)
        for (int ii = 0; ii < raw[k].m0(); ii++) at[raw[k].rp[ii]].push_back({(int) k, ii});
    std::vector<int> cpoff(raw.size());
    std::vector<int> cpool;
    for (size_t k = 0; k < raw.size(); k++)
    {
        cpoff[k] = (int) cpool.size();
        cpool.insert(cpool.end(), raw[k].cp.begin(), raw[k].cp.end());
    }
    std::vector<long long> obase, sstart;
    std::vector<int> sp{0}, slen, scp;
    long long off = 0, cells = 0;
    int longest = 1;
    for (size_t s = 0; s < co.out.size(); s++)
    {
        const Blk &o = co.out[s];
        for (int i = 0; i < o.m0(); i++)
        {
            obase.push_back(off + (long long) i * o.n0());
            // at[r] is in ascending block order and holds exactly the
            // signature's blocks (a block lists r at most once)
            for (const auto &kb : at[o.rp[i]])
            {
                const Blk &b = raw[kb.first];
                if (b.n0() == 0) continue;
                sstart.push_back(raw_off[kb.first] + (long long) kb.second * b.n0());
                slen.push_back(b.n0());
                scp.push_back(cpoff[kb.first]);
                cells += b.n0();
                longest = std::max(longest, b.n0());
            }
            sp.push_back((int) slen.size());
        }
        off += o.size();
    }
    P.nrows = (int) obase.size();
    P.ncolw = (n + 31) / 32 * 32;
    P.h = 1;
    while (256 * P.h < longest) P.h *= 2;
    P.cells = cells;
    if (P.nrows == 0) return;
    P.obase = upload(obase.data(), (int64_t) obase.size());
    P.sp = upload(sp.data(), (int64_t) sp.size());
    if (!slen.empty())
    {
        P.sstart = upload(sstart.data(), (int64_t) sstart.size());
        P.slen = upload(slen.data(), (int64_t) slen.size());
        P.scp = upload(scp.data(), (int64_t) scp.size());
    }
    if (cpool.empty()) cpool.push_back(0);
    P.cpool = upload(cpool.data(), (int64_t) cpool.size());
}

// -------------------------------------------------- structural operations ----
// global row of every value of a PD / SPD (for DA), built once
static const int *value_rows(matrix *A)
{
    DBuf<int> *slot = A->kind == MATRIX_PERMUTED_DENSE ? &as_pd(A)->drow : &as_spd(A)->drow;
    if (!slot->p)
    {
        std::vector<PDM *> bl;
        std::vector<long long> off;
        blocks_of(A, bl, off);
        std::vector<int> rows;
        for (PDM *b : bl)
            for (int ii = 0; ii < b->m0; ii++)
                for (int jj = 0; jj < b->n0; jj++) rows.push_back(b->row_perm[ii]);
        *slot = upload(rows.data(), (int64_t) rows.size());
        sync();
    }
    return slot->p;
}

namespace {
struct DAB : Bound
{
    DBuf<double> d;
};
} // namespace

matrix *pd_copy_sparsity(const matrix *A)
{
    std::vector<PDM *> bl;
    std::vector<long long> off;
    blocks_of(A, bl, off);
    std::vector<Blk> blocks;
    for (PDM *b : bl) blocks.push_back(blk_of(b));
    std::vector<int> src_p, src_idx;
    if (A->kind == MATRIX_STACKED_PD)
    {
        src_p = as_spd(A)->src_p;
        src_idx = as_spd(A)->src_idx;
        src_p.assign(blocks.size() + 1, 0);
        src_idx.assign(blocks.size(), 0);
        for (size_t k = 0; k < blocks.size(); k++)
        {
            src_p[k + 1] = (int) k + 1;
            src_idx[k] = (int) k;
        }
    }
    matrix *C = make_output_blocks(A->kind == MATRIX_STACKED_PD, A->m, A->n, blocks, src_p, src_idx);
    // bind what the DA fill needs (DA_matrices_fill_values reads C from here)
    value_rows(const_cast<matrix *>(A));
    auto b = std::make_unique<DAB>();
    b->d.alloc(std::max(A->m, 1));
    set_bound(C, std::move(b));
    return C;
}

void pd_DA_fill(const double *d, const matrix *A, matrix *C)
{
    // DA_pd_fill_values (permuted_dense_linalg.c:76-85): copy, then scale each
    // dense row by d[row_perm[ii]] -- one product per value
    matrix *Am = const_cast<matrix *>(A);
    DAB *b = bound_of<DAB>(C, "DA_matrices_fill_values");
    const int *rows = value_rows(Am);
    h2d(b->d.p, d, (size_t) A->m * sizeof(double));
    scale_rows(A->nnz, rows, mat_vals(Am), b->d.p, out_buf(C), stream());
    out_publish(C);
}

// transpose of PD (permuted_dense_linalg.c:35-46) and SPD (stacked_pd_linalg.c:66-95):
// swapped perms; the SPD's per-block transposes are coalesced (disjoint cells,
// plain copies), composed into one gather.
matrix *pd_transpose_alloc(const matrix *A)
{
    std::vector<PDM *> bl;
    std::vector<long long> off;
    blocks_of(A, bl, off);
    std::vector<Blk> raw;
    std::vector<std::vector<int>> raw_src;  // raw cell -> source position
    for (size_t k = 0; k < bl.size(); k++)
    {
        PDM *B = bl[k];
        Blk b;
        b.m = A->n;
        b.n = A->m;
        b.rp = B->col_perm;
        b.cp = B->row_perm;
        std::vector<int> src;
        for (int jj = 0; jj < B->n0; jj++)
            for (int ii = 0; ii < B->m0; ii++) src.push_back((int) (off[k] + (long long) ii * B->n0 + jj));
        raw.push_back(b);
        raw_src.push_back(std::move(src));
    }
    std::vector<int> perm;
    matrix *C;
    if (A->kind == MATRIX_PERMUTED_DENSE)
    {
        std::vector<int> src_p{0, 1}, src_idx{0};
        C = make_output_blocks(false, A->n, A->m, raw, src_p, src_idx);
        perm = raw_src[0];
    }
    else
    {
        long long raw_total = 0;
        std::vector<long long> raw_off = offsets_of(raw, &raw_total);
        Coalesced co = coalesce_blocks(A->n, A->m, raw, raw_off);
        C = make_output_blocks(true, A->n, A->m, co.out, co.src_p, co.src_idx);
        std::vector<int> flat;
        for (auto &v : raw_src) flat.insert(flat.end(), v.begin(), v.end());
        for (size_t e = 0; e + 1 < co.seg.size(); e++)
        {
            if (co.seg[e + 1] - co.seg[e] != 1) api_exit("transpose_matrices_alloc", "overlapping cells");
            perm.push_back(flat[co.idx[co.seg[e]]]);
        }
    }
    auto b = std::make_unique<GatherB>();
    b->perm = upload(perm.data(), (int64_t) perm.size());
    set_bound(C, std::move(b));
    return C;
}

void pd_transpose_fill(const matrix *A, matrix *AT)
{
    GatherB *b = bound_of<GatherB>(AT, "transpose_matrices_fill_values");
    gather_values(A->nnz, b->perm.p, mat_vals(const_cast<matrix *>(A)), out_buf(AT), stream());
    out_publish(AT);
}

} // namespace sdl

using namespace sdl;

extern "C" {
#pragma GCC visibility push(default)

matrix *new_permuted_dense(int m, int n, int m0, int n0, const int *row_perm, const int *col_perm,
                           const double *X_data)
{
    Blk b;
    b.m = m;
    b.n = n;
    if (m0 > 0) b.rp.assign(row_perm, row_perm + m0);
    if (n0 > 0) b.cp.assign(col_perm, col_perm + n0);
    return &new_pdm(b, X_data)->base;
}

matrix *new_permuted_dense_full(int m, int n, const double *data)
{
    Blk b;
    b.m = m;
    b.n = n;
    for (int i = 0; i < m; i++) b.rp.push_back(i);
    for (int j = 0; j < n; j++) b.cp.push_back(j);
    return &new_pdm(b, data)->base;
}

matrix *new_stacked_pd(int m, int n, int n_blocks, permuted_dense **blocks,
                       const int *src_block_idx_p, const int *src_block_idx)
{
    std::vector<PDM *> bl;
    for (int k = 0; k < n_blocks; k++) bl.push_back(reinterpret_cast<PDM *>(blocks[k]));
    // row_perms must be pairwise disjoint (stacked_pd.c:220-250)
    std::vector<char> seen(std::max(m, 1), 0);
    for (PDM *b : bl)
        for (int r : b->row_perm)
        {
            if (seen[r]) api_exit("new_stacked_pd", "block row_perms overlap");
            seen[r] = 1;
        }
    std::vector<int> sp, si;
    if (src_block_idx_p)
    {
        sp.assign(src_block_idx_p, src_block_idx_p + n_blocks + 1);
        si.assign(src_block_idx, src_block_idx + sp[n_blocks]);
    }
    else
    {
        for (int k = 0; k <= n_blocks; k++) sp.push_back(k);
        for (int k = 0; k < n_blocks; k++) si.push_back(k);
    }
    return &new_spdm(m, n, bl, sp, si)->base;
}

#pragma GCC visibility pop
} // extern "C"
