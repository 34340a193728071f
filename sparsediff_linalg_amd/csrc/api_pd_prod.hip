// Products involving permuted_dense / stacked_pd operands (reference:
// src/permuted_dense/permuted_dense_linalg.c, src/stacked_pd/stacked_pd_linalg.c,
// src/stacked_pd/stacked_pd_kron_linalg.c, routed as src/matmul_dispatchers.c).
//
// Each alloc computes the output structure with the reference's rules and
// emits a device program: grouped-GEMM tasks (dense x dense) and sparse x
// dense lists, writing either the output directly or per-block raw partials
// that are then coalesced in source-block order (memset + ordered += of
// stacked_pd_coalesce.c:326-381).  Association follows the reference per
// kind: B (d A) on the dense paths, (d A) B for pd x csc, B (d A) for csc x pd.
#include "pd_internal.h"

#include <algorithm>
#include <climits>
#include <cstring>

namespace sdl {

namespace {

// ------------------------------------------------------------ operands ----
// A dense block operand inside a value buffer (0: left operand B, 1: right A).
struct DOp
{
    const PDM *b;
    int buf;
    long long off;
};

std::vector<DOp> dops(const matrix *M, int buf)
{
    std::vector<DOp> v;
    if (M->kind == MATRIX_PERMUTED_DENSE) v.push_back({as_pd(M), buf, 0});
    else
    {
        SPDM *S = as_spd(M);
        for (size_t k = 0; k < S->blocks.size(); k++) v.push_back({S->blocks[k], buf, S->boff[k]});
    }
    return v;
}

// host copy of a sparse operand's CSC mirror structure (columns as rows; the
// value of entry q lives at CSR position cperm[q])
struct HostCsc
{
    std::vector<int> p, i, perm;
};

HostCsc host_csc(SparseM *S)
{
    sparse_ensure_csc(S);
    HostCsc h;
    const int n = S->base.n, nnz = S->base.nnz;
    h.p.resize((size_t) n + 1);
    h.i.resize(nnz);
    h.perm.resize(nnz);
    d2h(h.p.data(), S->cp.p, ((size_t) n + 1) * sizeof(int));
    if (nnz)
    {
        d2h(h.i.data(), S->ci.p, (size_t) nnz * sizeof(int));
        d2h(h.perm.data(), S->cperm.p, (size_t) nnz * sizeof(int));
    }
    return h;
}

// CSR arrays of a sparse operand, read as the CSC of its transpose
HostCsc host_csr_as_csc(SparseM *S)
{
    HostCsc h;
    const CSR_matrix *c = S->csr;
    h.p.assign(c->p, c->p + c->m + 1);
    h.i.assign(c->i, c->i + c->nnz);
    h.perm.resize(c->nnz);
    for (int t = 0; t < c->nnz; t++) h.perm[t] = t;
    return h;
}

std::vector<int> iota_vec(int n, int base = 0)
{
    std::vector<int> v(n);
    for (int i = 0; i < n; i++) v[i] = base + i;
    return v;
}

// ------------------------------------------------------------ program ----
struct SdmmJob
{
    SdmmPlan plan;
    int dense_buf;        // 0 / 1: which operand holds the dense block
    long long dense_off;
    long long out_off;
    bool sparse_is_B;     // the sparse operand is the left one
};

struct ProdB : Bound
{
    GemmPlan gemm;
    std::vector<std::unique_ptr<SdmmJob>> sdmm;
    bool to_raw = false;
    DBuf<double> raw;
    PullList coalesce;
    RowCoalesce rowc;              // used instead of the pull lists when by_row
    bool by_row = false;
    DBuf<double> d;
    int d_len = 0;
    matrix *left_view = nullptr;   // when the left operand is read through a sparse view
    long long raw_cells = 0;       // per-block partial cells before the coalesce
    long long coalesce_terms = 0;  // raw cells read by the coalesce
};

struct Program
{
    GemmBuilder gb;
    std::vector<std::unique_ptr<SdmmJob>> sdmm;
};

// ---- per-block emitters (each writes one dense output block at out_off) ----

// B_blk^T diag(d) A_blk over their common rows (BTA_pd_pd_core,
// permuted_dense_linalg.c:164-231); qmap maps output columns to A's local
// columns (-1 offset: identity).  Returns false if the rows do not overlap.
bool emit_bta(GemmBuilder &g, const DOp &B, const DOp &A, std::vector<int> &ib,
              std::vector<int> &ia, int qmap_off, int t_begin_unused = 0)
{
    (void) t_begin_unused;
    int s = sorted_intersect(B.b->row_perm, A.b->row_perm, ib, ia);
    if (s == 0) return false;
    std::vector<int> w(s);
    for (int t = 0; t < s; t++) w[t] = A.b->row_perm[ia[t]];
    GTask T;
    T.p_off = B.off;
    T.q_off = A.off;
    T.p_buf = B.buf;
    T.q_buf = A.buf;
    T.p_ld = B.b->n0;
    T.q_ld = A.b->n0;
    T.p_mode = 0;
    T.q_mode = 0;
    T.S = s;
    T.prow = g.add_ints(ib);
    T.pmap = -1;
    T.qrow = g.add_ints(ia);
    T.qmap = qmap_off;
    T.widx = g.add_ints(w);
    g.tasks.push_back(T);
    return true;
}

// B_blk A_blk over B's columns matching A's rows (BA_pd_pd_fill_values,
// permuted_dense_linalg.c:482-540): P(s, i) = B[i][ib[s]].
bool emit_ba(GemmBuilder &g, const DOp &B, const DOp &A, int qmap_off)
{
    std::vector<int> ib, ia;
    int s = sorted_intersect(B.b->col_perm, A.b->row_perm, ib, ia);
    if (s == 0) return false;
    GTask T;
    T.p_off = B.off;
    T.q_off = A.off;
    T.p_buf = B.buf;
    T.q_buf = A.buf;
    T.p_ld = B.b->n0;
    T.q_ld = A.b->n0;
    T.p_mode = 1;
    T.q_mode = 0;
    T.S = s;
    T.prow = g.add_ints(ib);
    T.pmap = -1;
    T.qrow = g.add_ints(ia);
    T.qmap = qmap_off;
    T.widx = -1;
    g.tasks.push_back(T);
    return true;
}

// Structure + tasks of B_blk^T D A where A is a list of blocks (pd or spd):
// output rows = B.col_perm, cols = sorted union of the contributing blocks'
// col_perms (BTA_pd_spd_alloc, stacked_pd_linalg.c:250-316).  With
// b_transposed the left operand is used as B^T (BA_pd_spd: P(s,i) = B[i][s]).
Blk bta_pd_blocks(GemmBuilder &g, const DOp &B, const std::vector<DOp> &As, long long out_off,
                  int Bm_out, int An, bool b_transposed)
{
    Blk out;
    out.m = Bm_out;
    out.n = An;
    out.rp = b_transposed ? B.b->row_perm : B.b->col_perm;
    const std::vector<int> &brows = b_transposed ? B.b->col_perm : B.b->row_perm;
    std::vector<const std::vector<int> *> cps;
    std::vector<size_t> contrib;
    for (size_t k = 0; k < As.size(); k++)
        if (has_overlap(brows, As[k].b->row_perm))
        {
            cps.push_back(&As[k].b->col_perm);
            contrib.push_back(k);
        }
    out.cp = sorted_union(cps);
    const int t0 = (int) g.tasks.size();
    for (size_t k : contrib)
    {
        const DOp &A = As[k];
        std::vector<int> qmap(out.cp.size());
        for (size_t j = 0; j < out.cp.size(); j++) qmap[j] = A.b->col_inv[out.cp[j]];
        int qm = g.add_ints(qmap);
        if (b_transposed)
        {
            std::vector<int> ib, ia;
            int s = sorted_intersect(B.b->col_perm, A.b->row_perm, ib, ia);
            GTask T;
            T.p_off = B.off;
            T.q_off = A.off;
            T.p_buf = B.buf;
            T.q_buf = A.buf;
            T.p_ld = B.b->n0;
            T.q_ld = A.b->n0;
            T.p_mode = 1;
            T.q_mode = 0;
            T.S = s;
            T.prow = g.add_ints(ib);
            T.pmap = -1;
            T.qrow = g.add_ints(ia);
            T.qmap = qm;
            T.widx = -1;
            g.tasks.push_back(T);
        }
        else
        {
            std::vector<int> ib, ia;
            emit_bta(g, B, A, ib, ia, qm);
        }
    }
    g.add_output(out_off, out.n0(), out.m0(), out.n0(), t0, (int) g.tasks.size());
    return out;
}

// B_blk^T D A with A sparse read by columns (BTA_pd_csc_alloc /
// BTDA_pd_csc_fill_values, permuted_dense_linalg.c:542-644):
// C[i][j] = sum over A's column entries (k ascending) of ((d[k] A[k,j]) B[kk][i]).
Blk bta_pd_csc(Program &pg, const DOp &B, const HostCsc &Ac, int An, long long out_off, int Bn)
{
    Blk out;
    out.m = Bn;
    out.n = An;
    out.rp = B.b->col_perm;
    std::vector<int> lp{0}, dr, vp, dk;
    for (int j = 0; j < An; j++)
    {
        bool hit = false;
        for (int q = Ac.p[j]; q < Ac.p[j + 1]; q++)
            if (B.b->row_inv[Ac.i[q]] >= 0)
            {
                hit = true;
                break;
            }
        if (!hit) continue;
        out.cp.push_back(j);
        for (int q = Ac.p[j]; q < Ac.p[j + 1]; q++)
        {
            int kk = B.b->row_inv[Ac.i[q]];
            if (kk < 0) continue;
            dr.push_back(kk);
            vp.push_back(Ac.perm[q]);
            dk.push_back(Ac.i[q]);
        }
        lp.push_back((int) dr.size());
    }
    auto job = std::make_unique<SdmmJob>();
    sdmm_build(job->plan, lp, dr, vp, dk);
    job->plan.NC = B.b->n0;
    job->plan.ld = B.b->n0;
    job->plan.dense_trans = false;
    job->plan.assoc = 0;
    job->plan.sL = 1;
    job->plan.sC = (long long) out.n0();
    job->dense_buf = B.buf;
    job->dense_off = B.off;
    job->out_off = out_off;
    job->sparse_is_B = false;
    pg.sdmm.push_back(std::move(job));
    return out;
}

// B^T D A_blk with B sparse read by columns (BTA_csc_pd_alloc /
// BTDA_csc_pd_fill_values, permuted_dense_linalg.c:646-741):
// C[i][j] = sum over B's column entries of (B[k,i] (d[k] A[kk][j])).
// sparse_is_B = the sparse operand is the left one.
Blk bta_csc_pd(Program &pg, const HostCsc &Bc, int Bn, const DOp &A, long long out_off,
               bool with_d)
{
    Blk out;
    out.m = Bn;
    out.n = A.b->base.n;
    out.cp = A.b->col_perm;
    std::vector<int> lp{0}, dr, vp, dk;
    for (int i = 0; i < Bn; i++)
    {
        bool hit = false;
        for (int q = Bc.p[i]; q < Bc.p[i + 1]; q++)
            if (A.b->row_inv[Bc.i[q]] >= 0)
            {
                hit = true;
                break;
            }
        if (!hit) continue;
        out.rp.push_back(i);
        for (int q = Bc.p[i]; q < Bc.p[i + 1]; q++)
        {
            int kk = A.b->row_inv[Bc.i[q]];
            if (kk < 0) continue;
            dr.push_back(kk);
            vp.push_back(Bc.perm[q]);
            dk.push_back(with_d ? Bc.i[q] : -1);
        }
        lp.push_back((int) dr.size());
    }
    auto job = std::make_unique<SdmmJob>();
    sdmm_build(job->plan, lp, dr, vp, with_d ? dk : std::vector<int>());
    job->plan.NC = A.b->n0;
    job->plan.ld = A.b->n0;
    job->plan.dense_trans = false;
    job->plan.assoc = 1;
    job->plan.sL = (long long) out.n0();
    job->plan.sC = 1;
    job->dense_buf = A.buf;
    job->dense_off = A.off;
    job->out_off = out_off;
    job->sparse_is_B = true;
    pg.sdmm.push_back(std::move(job));
    return out;
}

// B_blk S with S sparse CSR (BA_pd_csr_alloc / _fill_values,
// permuted_dense_linalg.c:313-430): C[i][c] = sum over the compact column c
// (B's columns r ascending, then S's row order) of B[i][r] S[r,c].
Blk ba_pd_csr(Program &pg, const DOp &B, const CSR_matrix *S, long long out_off,
              const std::vector<int> &b_rows, const std::vector<int> &b_cols, int Bm)
{
    Blk out;
    out.m = Bm;
    out.n = S->n;
    out.rp = b_rows;
    std::vector<int> mark(S->n, -1);
    for (int row : b_cols)
        for (int q = S->p[row]; q < S->p[row + 1]; q++) mark[S->i[q]] = 0;
    for (int j = 0; j < S->n; j++)
        if (mark[j] >= 0)
        {
            mark[j] = (int) out.cp.size();
            out.cp.push_back(j);
        }
    std::vector<std::vector<std::pair<int, int>>> col(out.cp.size());
    for (size_t r = 0; r < b_cols.size(); r++)
    {
        int row = b_cols[r];
        for (int q = S->p[row]; q < S->p[row + 1]; q++) col[mark[S->i[q]]].push_back({(int) r, q});
    }
    std::vector<int> lp{0}, dr, vp;
    for (auto &c : col)
    {
        for (auto &e : c)
        {
            dr.push_back(e.first);
            vp.push_back(e.second);
        }
        lp.push_back((int) dr.size());
    }
    auto job = std::make_unique<SdmmJob>();
    sdmm_build(job->plan, lp, dr, vp, std::vector<int>());
    job->plan.NC = (int) b_rows.size();
    job->plan.ld = (int) b_cols.size();
    job->plan.dense_trans = true;
    job->plan.assoc = 0;
    job->plan.sL = 1;
    job->plan.sC = (long long) out.n0();
    job->dense_buf = B.buf;
    job->dense_off = B.off;
    job->out_off = out_off;
    job->sparse_is_B = false;
    pg.sdmm.push_back(std::move(job));
    return out;
}

// ------------------------------------------------------------ assembly ----
// Output from one block (PD) or from per-block raw partials coalesced (SPD).
matrix *finish(Program &pg, bool stack_out, int m, int n, const std::vector<Blk> &raw,
               long long raw_total, const std::vector<long long> &raw_off, int d_len)
{
    auto b = std::make_unique<ProdB>();
    matrix *C;
    if (!stack_out)
    {
        std::vector<int> sp{0, 1}, si{0};
        Blk one = raw.empty() ? Blk{m, n, {}, {}} : raw[0];
        one.m = m;
        one.n = n;
        PDM *P = new_pdm(one, nullptr);
        C = &P->base;
        b->to_raw = false;
    }
    else
    {
        const bool by_row = rowc_eligible(m, n, raw);
        Coalesced co = coalesce_blocks(m, n, raw, raw_off, !by_row);
        std::vector<PDM *> pds;
        for (const Blk &o : co.out) pds.push_back(new_pdm(o, nullptr));
        C = &new_spdm(m, n, pds, co.src_p, co.src_idx)->base;
        b->to_raw = true;
        b->raw.alloc(raw_total);
        b->raw_cells = raw_total;
        b->by_row = by_row;
        if (by_row)
        {
            rowc_build(b->rowc, m, n, raw, raw_off, co);
            b->coalesce_terms = b->rowc.cells;
        }
        else
        {
            b->coalesce_terms = (long long) co.idx.size();
            pull_build(b->coalesce, co.seg, co.idx);
        }
    }
    gemm_upload(b->gemm, pg.gb);
    b->sdmm = std::move(pg.sdmm);
    b->d.alloc(std::max(d_len, 1));
    b->d_len = d_len;
    if (C->kind == MATRIX_PERMUTED_DENSE) as_pd(C)->bound = std::move(b);
    else as_spd(C)->bound = std::move(b);
    return C;
}

ProdB *prod_of(matrix *C, const char *fn)
{
    std::unique_ptr<Bound> &slot = C->kind == MATRIX_PERMUTED_DENSE ? as_pd(C)->bound : as_spd(C)->bound;
    ProdB *b = slot ? dynamic_cast<ProdB *>(slot.get()) : nullptr;
    if (!b) api_exit(fn, "output was not produced by the matching alloc");
    return b;
}

// The device part of a fill: every kernel of C's program on the library
// stream, d already in device memory (or null), C's device values written.
void launch(ProdB *b, matrix *B, matrix *A, const double *dd, matrix *C, hipEvent_t e0,
            hipEvent_t e1, hipEvent_t e_gemm = nullptr)
{
    const double *dB = mat_vals(b->left_view ? b->left_view : B);
    const double *dA = mat_vals(A);
    double *target = b->to_raw ? b->raw.p : out_buf(C);
    if (e0) SDL_HIP(hipEventRecord(e0, stream()));
    gemm_run(b->gemm, dB, dA, dd, target, stream());
    for (auto &j : b->sdmm)
    {
        const double *dense = (j->dense_buf == 0 ? dB : dA) + j->dense_off;
        const double *sv = j->sparse_is_B ? dB : dA;
        sdmm_run(j->plan, sv, dense, dd, target + j->out_off, stream());
    }
    if (e_gemm) SDL_HIP(hipEventRecord(e_gemm, stream()));
    if (b->to_raw)
    {
        if (b->by_row) rowc_run(b->rowc, b->raw.p, out_buf(C), stream());
        else pull_sum(b->coalesce, b->raw.p, out_buf(C), stream());
    }
    if (e1) SDL_HIP(hipEventRecord(e1, stream()));
}

// left = B (buf 0), right = A (buf 1); sparse operands are read through their
// CSR values (the lists hold CSR positions).
void run(const char *fn, matrix *B, matrix *A, const double *d, matrix *C)
{
    ProdB *b = prod_of(C, fn);
    const double *dd = nullptr;
    if (d && b->d_len > 0)
    {
        h2d(b->d.p, d, (size_t) b->d_len * sizeof(double));
        dd = b->d.p;
    }
    launch(b, B, A, dd, C, nullptr, nullptr);
    out_publish(C);
}

std::vector<long long> offsets(const std::vector<Blk> &bl, long long *total)
{
    std::vector<long long> off;
    long long t = 0;
    for (const Blk &b : bl)
    {
        off.push_back(t);
        t += b.size();
    }
    *total = t;
    return off;
}

// Emit per-block partials one after the other: `emit(k, out_off)` builds
// block k's partial and returns its structure.
template <typename F>
void partials(int nblocks, std::vector<Blk> &raw, std::vector<long long> &raw_off, long long &total,
              F &&emit)
{
    total = 0;
    for (int k = 0; k < nblocks; k++)
    {
        Blk b = emit(k, total);
        raw.push_back(b);
        raw_off.push_back(total);
        total += b.size();
    }
}

} // namespace

// ================================================================ B^T A ====
matrix *pd_BTA_alloc(matrix *A, matrix *B)
{
    if (A->m != B->m) api_exit("BTA_matrices_alloc", "A and B row counts differ");
    Program pg;
    std::vector<Blk> raw;
    std::vector<long long> raw_off;
    long long total = 0;
    if (B->kind == MATRIX_PERMUTED_DENSE)
    {
        DOp Bo = dops(B, 0)[0];
        if (A->kind == MATRIX_SPARSE)
        {
            HostCsc Ac = host_csc(as_sparse(A));
            raw.push_back(bta_pd_csc(pg, Bo, Ac, A->n, 0, B->n));
        }
        else if (A->kind == MATRIX_PERMUTED_DENSE && !has_overlap(Bo.b->row_perm, as_pd(A)->row_perm))
        {
            raw.push_back(Blk{B->n, A->n, {}, {}});  // empty (permuted_dense_linalg.c:121-124)
        }
        else
        {
            raw.push_back(bta_pd_blocks(pg.gb, Bo, dops(A, 1), 0, B->n, A->n, false));
        }
        return finish(pg, false, B->n, A->n, raw, raw[0].size(), {0}, A->m);
    }
    if (B->kind == MATRIX_STACKED_PD)
    {
        // blockwise over B's blocks, then coalesce (stacked_pd_linalg.c:117-155, 430-559)
        std::vector<DOp> Bs = dops(B, 0);
        HostCsc Ac;
        if (A->kind == MATRIX_SPARSE) Ac = host_csc(as_sparse(A));
        partials((int) Bs.size(), raw, raw_off, total, [&](int k, long long off) -> Blk {
            if (A->kind == MATRIX_SPARSE) return bta_pd_csc(pg, Bs[k], Ac, A->n, off, B->n);
            if (A->kind == MATRIX_PERMUTED_DENSE && !has_overlap(Bs[k].b->row_perm, as_pd(A)->row_perm))
                return Blk{B->n, A->n, {}, {}};
            return bta_pd_blocks(pg.gb, Bs[k], dops(A, 1), off, B->n, A->n, false);
        });
        return finish(pg, true, B->n, A->n, raw, total, raw_off, A->m);
    }
    // B sparse: A PD -> PD; A SPD -> blockwise over A's blocks (Cm = B->n)
    SparseM *SB = as_sparse(B);
    HostCsc Bc = host_csc(SB);
    std::vector<DOp> As = dops(A, 1);
    if (A->kind == MATRIX_PERMUTED_DENSE)
    {
        raw.push_back(bta_csc_pd(pg, Bc, B->n, As[0], 0, true));
        return finish(pg, false, B->n, A->n, raw, raw[0].size(), {0}, A->m);
    }
    partials((int) As.size(), raw, raw_off, total,
             [&](int k, long long off) -> Blk { return bta_csc_pd(pg, Bc, B->n, As[k], off, true); });
    return finish(pg, true, B->n, A->n, raw, total, raw_off, A->m);
}

void pd_BTDA_fill(matrix *A, const double *d, matrix *B, matrix *C)
{
    run("BTDA_matrices_fill_values", B, A, d, C);
}

// ============================================================== A^T D A ====
matrix *pd_ATA_alloc(matrix *A)
{
    Program pg;
    std::vector<DOp> As = dops(A, 1);
    std::vector<Blk> raw;
    std::vector<long long> raw_off;
    long long total = 0;
    auto emit = [&](int k, long long off) -> Blk {
        const DOp &a = As[k];
        Blk out{A->n, A->n, a.b->col_perm, a.b->col_perm};
        const int t0 = (int) pg.gb.tasks.size();
        GTask T;
        T.p_off = a.off;
        T.q_off = a.off;
        T.p_buf = 1;
        T.q_buf = 1;
        T.p_ld = a.b->n0;
        T.q_ld = a.b->n0;
        T.p_mode = 0;
        T.q_mode = 0;
        T.S = a.b->m0;
        std::vector<int> rows = iota_vec(a.b->m0);
        T.prow = pg.gb.add_ints(rows);
        T.qrow = T.prow;
        T.pmap = -1;
        T.qmap = -1;
        T.widx = pg.gb.add_ints(a.b->row_perm);
        pg.gb.tasks.push_back(T);
        pg.gb.add_output(off, out.n0(), out.m0(), out.n0(), t0, (int) pg.gb.tasks.size());
        return out;
    };
    if (A->kind == MATRIX_PERMUTED_DENSE)
    {
        raw.push_back(emit(0, 0));
        return finish(pg, false, A->n, A->n, raw, raw[0].size(), {0}, A->m);
    }
    partials((int) As.size(), raw, raw_off, total, emit);
    return finish(pg, true, A->n, A->n, raw, total, raw_off, A->m);
}

void pd_ATDA_fill(const matrix *A, const double *d, matrix *C)
{
    matrix *Am = const_cast<matrix *>(A);
    run("ATDA_matrices_fill_values", Am, Am, d, C);
}

// ================================================================= B A ====
matrix *pd_BA_alloc(const matrix *Bc, matrix *A)
{
    matrix *B = const_cast<matrix *>(Bc);
    if (B->n != A->m) api_exit("BA_matrices_alloc", "inner dimensions differ");
    Program pg;
    std::vector<Blk> raw;
    std::vector<long long> raw_off;
    long long total = 0;
    if (B->kind == MATRIX_PERMUTED_DENSE)
    {
        DOp Bo = dops(B, 0)[0];
        if (A->kind == MATRIX_SPARSE)
        {
            sparse_ensure_dev(as_sparse(A));  // the fill reads A's device values
            raw.push_back(ba_pd_csr(pg, Bo, as_sparse(A)->csr, 0, Bo.b->row_perm, Bo.b->col_perm, B->m));
        }
        else if (A->kind == MATRIX_PERMUTED_DENSE)
        {
            DOp Ao = dops(A, 1)[0];
            if (!has_overlap(Bo.b->col_perm, Ao.b->row_perm)) raw.push_back(Blk{B->m, A->n, {}, {}});
            else
            {
                Blk out{B->m, A->n, Bo.b->row_perm, Ao.b->col_perm};
                const int t0 = (int) pg.gb.tasks.size();
                emit_ba(pg.gb, Bo, Ao, -1);
                pg.gb.add_output(0, out.n0(), out.m0(), out.n0(), t0, (int) pg.gb.tasks.size());
                raw.push_back(out);
            }
        }
        else
            raw.push_back(bta_pd_blocks(pg.gb, Bo, dops(A, 1), 0, B->m, A->n, true));
        return finish(pg, false, B->m, A->n, raw, raw[0].size(), {0}, 0);
    }
    // B sparse: C = B A = BTA_csc_pd(B^T, A) with B's CSR read as the CSC of B^T
    // (matmul_dispatchers.c:147-194)
    SparseM *SB = as_sparse(B);
    sparse_ensure_dev(SB);
    HostCsc Bt = host_csr_as_csc(SB);
    std::vector<DOp> As = dops(A, 1);
    if (A->kind == MATRIX_PERMUTED_DENSE)
    {
        raw.push_back(bta_csc_pd(pg, Bt, B->m, As[0], 0, false));
        return finish(pg, false, B->m, A->n, raw, raw[0].size(), {0}, 0);
    }
    partials((int) As.size(), raw, raw_off, total,
             [&](int k, long long off) -> Blk { return bta_csc_pd(pg, Bt, B->m, As[k], off, false); });
    return finish(pg, true, B->m, A->n, raw, total, raw_off, 0);
}

void pd_BA_fill(const matrix *B, const matrix *A, matrix *C)
{
    run("BA_matrices_fill_values", const_cast<matrix *>(B), const_cast<matrix *>(A), nullptr, C);
}

// ================================================================ kron ====
// y = kron(I_p, A) x for a full PD A: y (p x m) = x (p x n) A^T, one GEMM
// (permuted_dense.c:131-150).
void pd_kron_vec(const matrix *Am, const double *x, double *y, int p)
{
    PDM *A = as_pd(Am);
    if (A->m0 != Am->m || A->n0 != Am->n) api_exit("BA_kron_vec_matrices", "A must be a full permuted_dense");
    const int m = Am->m, n = Am->n;
    if (p <= 0 || m == 0) return;
    DBuf<double> dx = upload(x, (int64_t) p * n);
    DBuf<double> dy((size_t) p * m);
    if (n == 0)
    {
        dzero(dy.p, dy.bytes());
    }
    else
    {
        GemmBuilder g;
        GTask T;
        T.p_off = 0;
        T.q_off = 0;
        T.p_buf = 0;  // x
        T.q_buf = 1;  // A
        T.p_ld = n;
        T.q_ld = n;
        T.p_mode = 1;  // P(s=j, i=b) = x[b*n + j]
        T.q_mode = 1;  // Q(s=j, c=i) = A[i*n + j]
        T.S = n;
        T.prow = g.add_ints(iota_vec(n));
        T.pmap = -1;
        T.qrow = T.prow;
        T.qmap = -1;
        T.widx = -1;
        g.tasks.push_back(T);
        g.add_output(0, m, p, m, 0, 1);
        GemmPlan gp;
        gemm_upload(gp, g);
        gemm_run(gp, dx.p, pd_vals(A), nullptr, dy.p, stream());
    }
    d2h(y, dy.p, (size_t) p * m * sizeof(double));
}

// C = kron(I_p, A) J (I_A) or kron(A, I_p) J (A_I) for a full PD A
// (stacked_pd_kron_linalg.c:48-307): block k of the Kronecker factor has rows
// R_k and columns K_k over A's dense values; per block the product is
// BA_pd_pd / BA_pd_csr / BTA_pd_spd(block^T, J), empty blocks dropped.
matrix *pd_dense_kron_alloc(const matrix *Am, int p, matrix *J, bool kron_A_I)
{
    PDM *A = as_pd(Am);
    if (A->m0 != Am->m || A->n0 != Am->n) api_exit("BA_dense_kron", "A must be a full permuted_dense");
    const int m = A->m0, n = A->n0;
    if (J->m != p * n) api_exit("BA_dense_kron", "J->m must equal p * A->n");
    Program pg;
    std::vector<Blk> outs;
    std::vector<int> src_idx;
    long long total = 0;
    std::vector<long long> out_off;
    const CSR_matrix *Jcsr = J->kind == MATRIX_SPARSE ? as_sparse(J)->csr : nullptr;
    if (Jcsr) sparse_ensure_dev(as_sparse(J));
    std::vector<DOp> Js = J->kind == MATRIX_SPARSE ? std::vector<DOp>() : dops(J, 1);
    for (int k = 0; k < p; k++)
    {
        // the scratch block view of factor block k, over A's values (buffer 0)
        PDM blk;
        blk.m0 = m;
        blk.n0 = n;
        for (int i = 0; i < m; i++) blk.row_perm.push_back(kron_A_I ? k + p * i : k * m + i);
        for (int j = 0; j < n; j++) blk.col_perm.push_back(kron_A_I ? k + p * j : k * n + j);
        blk.base.m = m * p;
        blk.base.n = n * p;
        blk.row_inv.assign(m * p, -1);
        blk.col_inv.assign(n * p, -1);
        for (int i = 0; i < m; i++) blk.row_inv[blk.row_perm[i]] = i;
        for (int j = 0; j < n; j++) blk.col_inv[blk.col_perm[j]] = j;
        DOp Bo{&blk, 0, 0};
        Blk out;
        const size_t tb = pg.gb.tasks.size(), jb = pg.gb.jobs.size(), sb = pg.sdmm.size(),
                     pb = pg.gb.pool.size();
        const long long fb = pg.gb.flops;
        if (Jcsr)
            out = ba_pd_csr(pg, Bo, Jcsr, total, blk.row_perm, blk.col_perm, m * p);
        else if (J->kind == MATRIX_PERMUTED_DENSE)
        {
            const DOp &Jo = Js[0];
            if (!has_overlap(blk.col_perm, Jo.b->row_perm)) out = Blk{m * p, J->n, {}, {}};
            else
            {
                out = Blk{m * p, J->n, blk.row_perm, Jo.b->col_perm};
                const int t0 = (int) pg.gb.tasks.size();
                emit_ba(pg.gb, Bo, Jo, -1);
                pg.gb.add_output(total, out.n0(), out.m0(), out.n0(), t0, (int) pg.gb.tasks.size());
            }
        }
        else
            out = bta_pd_blocks(pg.gb, Bo, Js, total, m * p, J->n, true);
        out.m = m * p;
        out.n = J->n;
        if (out.size() == 0)
        {
            // drop the empty block and anything it emitted
            pg.gb.tasks.resize(tb);
            pg.gb.jobs.resize(jb);
            pg.gb.pool.resize(pb);
            pg.gb.flops = fb;
            pg.sdmm.resize(sb);
            continue;
        }
        outs.push_back(out);
        out_off.push_back(total);
        src_idx.push_back(k);
        total += out.size();
    }
    // blocks are row-disjoint: the output is the stack of them, written directly
    auto b = std::make_unique<ProdB>();
    std::vector<PDM *> pds;
    for (const Blk &o : outs) pds.push_back(new_pdm(o, nullptr));
    std::vector<int> src_p(outs.size() + 1);
    for (size_t k = 0; k <= outs.size(); k++) src_p[k] = (int) k;
    SPDM *C = new_spdm(m * p, J->n, pds, src_p, src_idx);
    gemm_upload(b->gemm, pg.gb);
    b->sdmm = std::move(pg.sdmm);
    b->to_raw = false;
    b->d_len = 0;
    C->bound = std::move(b);
    return &C->base;
}

void pd_dense_kron_fill(const matrix *A, int p, matrix *J, matrix *C, bool kron_A_I)
{
    (void) p;
    (void) kron_A_I;
    run("BA_dense_kron_fill_values", const_cast<matrix *>(A), J, nullptr, C);
}

} // namespace sdl

// ====================================================== device refills ====
extern "C" {
#pragma GCC visibility push(default)

int sdl_block_refill_device_split(matrix *C, matrix *left, matrix *right, const double *dd,
                                  void *ev_start, void *ev_products, void *ev_stop)
{
    using namespace sdl;
    if (!C || !left || !right || (C->kind != MATRIX_PERMUTED_DENSE && C->kind != MATRIX_STACKED_PD))
    {
        set_error("sdl_block_refill_device: C must be a permuted_dense / stacked_pd product output");
        return 1;
    }
    std::unique_ptr<Bound> &slot = C->kind == MATRIX_PERMUTED_DENSE ? as_pd(C)->bound : as_spd(C)->bound;
    ProdB *b = slot ? dynamic_cast<ProdB *>(slot.get()) : nullptr;
    if (!b)
    {
        set_error("sdl_block_refill_device: C was not produced by a product alloc");
        return 1;
    }
    launch(b, left, right, b->d_len > 0 ? dd : nullptr, C, (hipEvent_t) ev_start,
           (hipEvent_t) ev_stop, (hipEvent_t) ev_products);
    return 0;
}

int sdl_block_refill_device(matrix *C, matrix *left, matrix *right, const double *dd,
                            void *ev_start, void *ev_stop)
{
    return sdl_block_refill_device_split(C, left, right, dd, ev_start, nullptr, ev_stop);
}

int sdl_block_stats(const matrix *C, double *out)
{
    using namespace sdl;
    if (!C || !out || (C->kind != MATRIX_PERMUTED_DENSE && C->kind != MATRIX_STACKED_PD)) return 1;
    const std::unique_ptr<Bound> &slot =
        C->kind == MATRIX_PERMUTED_DENSE ? as_pd(C)->bound : as_spd(C)->bound;
    const ProdB *b = slot ? dynamic_cast<const ProdB *>(slot.get()) : nullptr;
    if (!b) return 1;
    long long sflops = 0;
    for (const auto &j : b->sdmm) sflops += 2ll * j->plan.nent * j->plan.NC;
    out[0] = (double) b->gemm.flops;
    out[1] = (double) sflops;
    out[2] = (double) b->gemm.njobs;
    out[3] = (double) (b->gemm.tasks.n);
    out[4] = (double) b->raw_cells;
    out[5] = (double) C->nnz;
    out[6] = (double) b->coalesce_terms;
    out[7] = (double) b->sdmm.size();
    return 0;
}

double *sdl_block_device_values(matrix *C)
{
    using namespace sdl;
    if (!C || (C->kind != MATRIX_PERMUTED_DENSE && C->kind != MATRIX_STACKED_PD)) return nullptr;
    return out_buf(C);
}

int sdl_block_publish(matrix *C)
{
    using namespace sdl;
    if (!C || (C->kind != MATRIX_PERMUTED_DENSE && C->kind != MATRIX_STACKED_PD)) return 1;
    out_publish(C);
    return 0;
}

#pragma GCC visibility pop
} // extern "C"
