/*
 * ORACLE — test infrastructure only.
 *
 * A single-threaded CPU restatement of the sparsediff_linalg sparse-product
 * algorithms, written from the reference's documented semantics.  Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker (or the timed CPU baseline), never as a product path.
 *
 * Every function cites the reference file:line it restates.  Offsets of the
 * product output are int64 so the same code also covers outputs past INT_MAX
 * (sampled-row checks); the reference itself stores them as int.
 *
 * Pinning status of each function (see tests/test_oracle.py):
 *   orc_spgemm_*          pinned: reference literals + bit-exact vs the
 *                         reference's own gustavson.c built in oracle/_ref +
 *                         brute-force dense boolean/FP64 products on tiny inputs
 *   orc_spgemm_range      a loop of orc_spgemm_row; equal to orc_spgemm_* on
 *                         random products (tests/test_oracle.py)
 *   orc_csr_transpose     pinned: literals + oracle/_ref + double-transpose identity
 *   orc_row_reduce_*      pinned: literals (test_row_reduce.h) + oracle/_ref vtable
 *   orc_coo_coalesce      pinned: equals orc_row_reduce on one-entry rows +
 *                         dense scatter-add closed form on small inputs
 *   orc_row_gather, orc_da, orc_kron_spmv, orc_coo_lower:
 *                         pinned: oracle/_ref + closed forms
 *   orc_pd_* / orc_spd_*  pinned against dense triple-loop products only
 *                         (the reference's CBLAS path is unbuildable here)
 */
#ifndef SDL_ORACLE_H
#define SDL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Gustavson C = kron(I_blocks, L) . diag(d) . A (src/sparse/gustavson.c) ---- */

/* Symbolic: gustavson.c:26-105.  Fills Cp[0..m_out] (m_out = blocks*Lm); if Ci
   is non-NULL also writes every row's columns sorted ascending.  Returns nnz. */
int64_t orc_spgemm_symbolic(int Lm, const int *Lp, const int *Li, int blocks,
                            int Am, int An, const int *Ap, const int *Ai,
                            int64_t *Cp, int *Ci);

/* Numeric, accumulator variant: gustavson.c:107-143.  acc has length An. */
void orc_spgemm_numeric(int Lm, const int *Lp, const int *Li, const double *Lx,
                        int blocks, const double *d, int Am, int An, const int *Ap,
                        const int *Ai, const double *Ax, const int64_t *Cp,
                        const int *Ci, double *acc, double *Cx);

/* Term count and plan: gustavson.c:145-205.  Returns the number of terms, or
   -1 when it exceeds max_terms (the reference then returns NULL).  plan may be
   NULL to only count. */
int64_t orc_spgemm_plan(int Lm, const int *Lp, const int *Li, int blocks, int Am,
                        int An, const int *Ap, const int *Ai, const int64_t *Cp,
                        const int *Ci, int64_t max_terms, int64_t *plan);

/* Numeric through a plan: gustavson.c:207-231. */
void orc_spgemm_numeric_planned(int Lm, const int *Lp, const int *Li,
                                const double *Lx, int blocks, const double *d,
                                int Am, const int *Ap, const double *Ax,
                                const int64_t *plan, int64_t nnzC, double *Cx);

/* One output row (sampled-row checks of products too large to form whole).
   mark (length An, initialised to -1 by the caller once) and acc (length An)
   are workspaces.  Returns the row's nnz; cols sorted, vals in that order. */
int orc_spgemm_row(int row, int Lm, const int *Lp, const int *Li, const double *Lx,
                   int blocks, const double *d, int Am, int An, const int *Ap,
                   const int *Ai, const double *Ax, int *mark, double *acc,
                   int *cols, double *vals);

/* The output rows [r0, r1) one after another through orc_spgemm_row (for
   block-wise checks of large products): Cp gets r1 - r0 + 1 offsets from 0,
   Ci / Cx the rows back to back.  Returns the range's nnz, or -1 if it would
   exceed cap (nothing beyond cap is written). */
int64_t orc_spgemm_range(int r0, int r1, int Lm, const int *Lp, const int *Li,
                         const double *Lx, int blocks, const double *d, int Am, int An,
                         const int *Ap, const int *Ai, const double *Ax, int *mark,
                         double *acc, int64_t cap, int64_t *Cp, int *Ci, double *Cx);

/* ---- transposition (src/sparse/CSC_matrix.c:62-189, CSR_matrix.c:99-205) ---- */

/* Stable counting-sort transpose of an m x n compressed matrix into n x m.
   Ax / ATx may be NULL (structure only).  iwork has length n. */
void orc_csr_transpose(int m, int n, const int *Ap, const int *Ai, const double *Ax,
                       int *ATp, int *ATi, double *ATx, int *iwork);

/* ---- row_reduce (src/sparse/sparse_matrix.c:136-240, CSR_sum.c:150-157) ---- */

/* Output structure and the source-nnz -> output-position map.  Cp has m_out+1
   entries; Ci and map need capacity A.nnz.  Returns output nnz. */
int orc_row_reduce_alloc(int m, int n, const int *Ap, const int *Ai,
                         const int *group, int m_out, int *Cp, int *Ci, int *map);
void orc_row_reduce_fill(int nnzA, const double *Ax, const int *map, int nnzC,
                         double *Cx);

/* ---- COO coalesce: row_reduce semantics with one triplet per source row ---- */

/* rows/cols/x of length nnz; Cp (m+1), Ci / Cx capacity nnz.  Duplicates are
   summed in input order starting from +0.0; columns ascending unique.
   Returns the output nnz. */
int64_t orc_coo_coalesce(int64_t nnz, const int *rows, const int *cols,
                         const double *x, int m, int n, int64_t *Cp, int *Ci,
                         double *Cx);

/* ---- small kernels ---- */

/* C = A[map, :] : sparse_matrix.c:79-134.  Cp has m_out+1 entries. */
int64_t orc_row_gather(int m_out, const int *map, const int *Ap, const int *Ai,
                       const double *Ax, int *Cp, int *Ci, double *Cx);

/* C = diag(d) A : CSR_matrix.c:86-97. */
void orc_da(int m, const int *Ap, const double *Ax, const double *d, double *Cx);

/* y = kron(I_p, A) x : linalg_sparse_matmuls.c:25-50. */
void orc_kron_spmv(int m, int n, const int *Ap, const int *Ai, const double *Ax,
                   const double *x, double *y, int p);

/* Lower-triangular COO export with value_map: COO_matrix.c:48-98.  Returns the
   count; rows/cols/x/value_map need capacity A.nnz. */
int orc_coo_lower(int m, const int *Ap, const int *Ai, const double *Ax, int *rows,
                  int *cols, double *x, int *value_map);

/* ---- matrix sums (src/sparse/CSR_sum.c:26-297, src/sparse/matrix_sum.c) ---- */

/* Union structure of nops operands of one shape.  P holds nops row-pointer
   arrays of m+1 entries each; I the operands' column indices back to back,
   operand k starting at off[k] (off has nops+1 entries).  Cp (m+1), Ci
   (capacity off[nops]) and map (off[nops]: output position of each operand
   entry).  Returns the output nnz. */
int orc_sum_n_alloc(int m, int n, int nops, const int *P, const int *I,
                    const int *off, int *Cp, int *Ci, int *map);
/* Values: +0.0 plus each operand's entries in operand order; d (nops x m row
   scales, or NULL) multiplies every term first. */
void orc_sum_n_fill(int m, int nops, const int *P, const int *off, const double *X,
                    const double *d, const int *map, int nnzC, double *Cx);

/* ---- dense Kronecker-vector products (src/support/mini_numpy.c:74-105) ---- */
void orc_y_kron_i_vec(int m, int k, int n, const double *Y, const double *w, double *v);
void orc_i_kron_xt_vec(int m, int k, int n, const double *X, const double *w, double *v);

#ifdef __cplusplus
}
#endif

#endif /* SDL_ORACLE_H */
