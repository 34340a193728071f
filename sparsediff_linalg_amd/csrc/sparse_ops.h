// Device sparse-format operations: stable transposition, value gathers,
// coalescing (row_reduce / COO), row gathers, diagonal scaling, kron SpMV.
#pragma once

#include "sdl_internal.h"

namespace sdl {

// Row index of every entry: rowidx[t] = r for Ap[r] <= t < Ap[r+1].
void expand_rows(int m, const int *dAp, int *drowidx, hipStream_t s);

// Stable transposition of an m x n compressed structure (CSR -> CSC, or CSC ->
// CSR): ATp[n+1], ATi[nnz] (source row of each entry, ascending within each
// output row) and perm[nnz] (source position of each output entry).  The
// reference's counting sort (CSC_matrix.c:62-104) gives the same order.
void transpose_structure(int m, int n, int nnz, const int *dAp, const int *dAi, int *dATp,
                         int *dATi, int *dperm);

// out[e] = in[perm[e]] for e < n.
void gather_values(int64_t n, const int *dperm, const double *din, double *dout,
                   hipStream_t s);

// Coalescing plan shared by row_reduce and COO -> CSR: items t = 0..n-1 with
// output row orow[t] and column ocol[t]; output entries are the distinct
// (row, col) pairs sorted; item order inside each entry is ascending t.
struct CoalescePlan
{
    int m_out = 0, n_out = 0;
    int64_t n_items = 0, nnz = 0;
    DBuf<int64_t> Cp;      // m_out + 1
    DBuf<int> Ci;          // nnz
    DBuf<int64_t> seg;     // nnz + 1: items of entry e are sorted[seg[e] .. seg[e+1])
    DBuf<unsigned> order;  // n_items: item index in sorted order
};
void coalesce_plan(CoalescePlan &P, int64_t n_items, const int *dorow, const int *docol,
                   int m_out, int n_out);
// Cx[e] = ((+0.0 + x[t1]) + x[t2]) + ... over entry e's items in ascending t.
void coalesce_fill(const CoalescePlan &P, const double *dx, double *dCx, hipStream_t s);

// C = A[map, :]: Cp from the gathered row lengths (host returns nnz);
// src[e] = A position feeding C entry e.
int64_t row_gather_plan(int m_out, const int *dmap, const int *dAp, int64_t *dCp64);
void row_gather_sources(int m_out, const int *dmap, const int *dAp, const int64_t *dCp64,
                        int64_t nnz, int *dsrc);

// C.x[t] = A.x[t] * d[rowidx[t]] (CSR_matrix.c:86-97 rounding: one product).
void scale_rows(int64_t nnz, const int *drowidx, const double *dAx, const double *dd,
                double *dCx, hipStream_t s);

// y_b = A x_b for b < p, each row a left-to-right dot (linalg_sparse_matmuls.c:25-50).
void kron_spmv(int m, int n, const int *dAp, const int *dAi, const double *dAx,
               const double *dx, double *dy, int p, hipStream_t s);

// int64 -> int32 narrowing copy (caller checked the range).
void narrow_i64(int64_t n, const int64_t *din, int *dout, hipStream_t s);

} // namespace sdl
