/*
 * ORACLE — test infrastructure only (see sdl_oracle.h).  CPU restatement of
 * the reference's sparse-product algorithms; each function cites the
 * reference file:line whose behaviour it follows.  Built with
 * -ffp-contract=off so every product and sum is individually rounded, as on
 * the reference's baseline x86-64 build.
 */
#include "sdl_oracle.h"

#include <stdlib.h>
#include <string.h>

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *) a, y = *(const int *) b;
    return (x > y) - (x < y);
}

/* gustavson.c:26-105: pass 1 stamps mark[j] = i to count distinct columns of
   row i; pass 2 stamps m_out + i and writes columns in first-seen order, then
   sorts the row. */
int64_t orc_spgemm_symbolic(int Lm, const int *Lp, const int *Li, int blocks,
                            int Am, int An, const int *Ap, const int *Ai,
                            int64_t *Cp, int *Ci)
{
    int64_t m_out = (int64_t) Lm * blocks;
    int nL = Am / blocks;
    int64_t *stamp = (int64_t *) malloc((size_t) (An > 0 ? An : 1) * sizeof(int64_t));
    for (int j = 0; j < An; j++) stamp[j] = -1;

    Cp[0] = 0;
    for (int64_t row = 0; row < m_out; row++)
    {
        int b = (int) (row / Lm), r = (int) (row % Lm);
        int64_t cnt = 0;
        for (int q = Lp[r]; q < Lp[r + 1]; q++)
        {
            int k = Li[q] + b * nL;
            for (int t = Ap[k]; t < Ap[k + 1]; t++)
            {
                int j = Ai[t];
                if (stamp[j] != row)
                {
                    stamp[j] = row;
                    cnt++;
                }
            }
        }
        Cp[row + 1] = Cp[row] + cnt;
    }
    if (Ci != NULL)
    {
        for (int64_t row = 0; row < m_out; row++)
        {
            int b = (int) (row / Lm), r = (int) (row % Lm);
            int64_t w = Cp[row];
            for (int q = Lp[r]; q < Lp[r + 1]; q++)
            {
                int k = Li[q] + b * nL;
                for (int t = Ap[k]; t < Ap[k + 1]; t++)
                {
                    int j = Ai[t];
                    if (stamp[j] != m_out + row)
                    {
                        stamp[j] = m_out + row;
                        Ci[w++] = j;
                    }
                }
            }
            qsort(Ci + Cp[row], (size_t) (Cp[row + 1] - Cp[row]), sizeof(int), cmp_int);
        }
    }
    free(stamp);
    return Cp[m_out];
}

/* gustavson.c:107-143: zero acc over C's row pattern, scatter
   acc[j] += (L.x * d[k]) * A.x in L-row order, gather back into C. */
void orc_spgemm_numeric(int Lm, const int *Lp, const int *Li, const double *Lx,
                        int blocks, const double *d, int Am, int An, const int *Ap,
                        const int *Ai, const double *Ax, const int64_t *Cp,
                        const int *Ci, double *acc, double *Cx)
{
    (void) An;
    int nL = Am / blocks;
    int64_t m_out = (int64_t) Lm * blocks;
    for (int64_t row = 0; row < m_out; row++)
    {
        int b = (int) (row / Lm), r = (int) (row % Lm);
        for (int64_t e = Cp[row]; e < Cp[row + 1]; e++) acc[Ci[e]] = 0.0;
        for (int q = Lp[r]; q < Lp[r + 1]; q++)
        {
            int k = Li[q] + b * nL;
            double v = Lx[q];
            if (d != NULL) v = v * d[k];
            for (int t = Ap[k]; t < Ap[k + 1]; t++) acc[Ai[t]] += v * Ax[t];
        }
        for (int64_t e = Cp[row]; e < Cp[row + 1]; e++) Cx[e] = acc[Ci[e]];
    }
}

/* gustavson.c:145-205: count terms with early exit against max_terms, then
   plan[t] = position in C.x of term t (pos[] filled from C's row). */
int64_t orc_spgemm_plan(int Lm, const int *Lp, const int *Li, int blocks, int Am,
                        int An, const int *Ap, const int *Ai, const int64_t *Cp,
                        const int *Ci, int64_t max_terms, int64_t *plan)
{
    int nL = Am / blocks;
    int64_t m_out = (int64_t) Lm * blocks;
    int64_t terms = 0;
    for (int64_t row = 0; row < m_out; row++)
    {
        int b = (int) (row / Lm), r = (int) (row % Lm);
        for (int q = Lp[r]; q < Lp[r + 1]; q++)
        {
            int k = Li[q] + b * nL;
            terms += Ap[k + 1] - Ap[k];
        }
        if (max_terms >= 0 && terms > max_terms) return -1;
    }
    if (plan == NULL || terms == 0) return terms;

    int64_t *pos = (int64_t *) malloc((size_t) (An > 0 ? An : 1) * sizeof(int64_t));
    int64_t t_out = 0;
    for (int64_t row = 0; row < m_out; row++)
    {
        int b = (int) (row / Lm), r = (int) (row % Lm);
        for (int64_t e = Cp[row]; e < Cp[row + 1]; e++) pos[Ci[e]] = e;
        for (int q = Lp[r]; q < Lp[r + 1]; q++)
        {
            int k = Li[q] + b * nL;
            for (int t = Ap[k]; t < Ap[k + 1]; t++) plan[t_out++] = pos[Ai[t]];
        }
    }
    free(pos);
    return t_out;
}

/* gustavson.c:207-231: memset C.x, then C.x[plan[t++]] += (L.x * d[k]) * A.x in
   the same term order as the accumulator fill. */
void orc_spgemm_numeric_planned(int Lm, const int *Lp, const int *Li,
                                const double *Lx, int blocks, const double *d,
                                int Am, const int *Ap, const double *Ax,
                                const int64_t *plan, int64_t nnzC, double *Cx)
{
    int nL = Am / blocks;
    memset(Cx, 0, (size_t) nnzC * sizeof(double));
    int64_t t_out = 0;
    for (int b = 0; b < blocks; b++)
    {
        for (int r = 0; r < Lm; r++)
        {
            for (int q = Lp[r]; q < Lp[r + 1]; q++)
            {
                int k = Li[q] + b * nL;
                double v = Lx[q];
                if (d != NULL) v = v * d[k];
                for (int t = Ap[k]; t < Ap[k + 1]; t++) Cx[plan[t_out++]] += v * Ax[t];
            }
        }
    }
}

/* One row of gustavson_alloc + gustavson_fill_values (gustavson.c:43-68,
   :78-101, :117-141).  mark must hold values != row on entry. */
int orc_spgemm_row(int row, int Lm, const int *Lp, const int *Li, const double *Lx,
                   int blocks, const double *d, int Am, int An, const int *Ap,
                   const int *Ai, const double *Ax, int *mark, double *acc,
                   int *cols, double *vals)
{
    (void) An;
    int nL = Am / blocks;
    int b = row / Lm, r = row % Lm;
    int cnt = 0;
    for (int q = Lp[r]; q < Lp[r + 1]; q++)
    {
        int k = Li[q] + b * nL;
        for (int t = Ap[k]; t < Ap[k + 1]; t++)
        {
            int j = Ai[t];
            if (mark[j] != row)
            {
                mark[j] = row;
                cols[cnt++] = j;
            }
        }
    }
    qsort(cols, (size_t) cnt, sizeof(int), cmp_int);
    for (int e = 0; e < cnt; e++) acc[cols[e]] = 0.0;
    for (int q = Lp[r]; q < Lp[r + 1]; q++)
    {
        int k = Li[q] + b * nL;
        double v = Lx[q];
        if (d != NULL) v = v * d[k];
        for (int t = Ap[k]; t < Ap[k + 1]; t++) acc[Ai[t]] += v * Ax[t];
    }
    for (int e = 0; e < cnt; e++) vals[e] = acc[cols[e]];
    return cnt;
}

int64_t orc_spgemm_range(int r0, int r1, int Lm, const int *Lp, const int *Li,
                         const double *Lx, int blocks, const double *d, int Am, int An,
                         const int *Ap, const int *Ai, const double *Ax, int *mark,
                         double *acc, int64_t cap, int64_t *Cp, int *Ci, double *Cx)
{
    int nL = Am / blocks;
    int64_t off = 0;
    Cp[0] = 0;
    for (int row = r0; row < r1; row++)
    {
        /* the row's term count bounds its nnz: never write past cap */
        int b = row / Lm, r = row % Lm;
        int64_t bound = 0;
        for (int q = Lp[r]; q < Lp[r + 1]; q++)
        {
            int k = Li[q] + b * nL;
            bound += Ap[k + 1] - Ap[k];
        }
        if (bound > An) bound = An;
        if (off + bound > cap) return -1;
        int cnt = orc_spgemm_row(row, Lm, Lp, Li, Lx, blocks, d, Am, An, Ap, Ai, Ax, mark, acc,
                                 Ci + off, Cx + off);
        off += cnt;
        Cp[row - r0 + 1] = off;
    }
    return off;
}

/* CSC_matrix.c:62-123 / CSR_matrix.c:99-142: column counts, exclusive scan,
   then a scatter in row order, so each output row lists sources ascending. */
void orc_csr_transpose(int m, int n, const int *Ap, const int *Ai, const double *Ax,
                       int *ATp, int *ATi, double *ATx, int *iwork)
{
    memset(iwork, 0, (size_t) n * sizeof(int));
    for (int t = 0; t < Ap[m]; t++) iwork[Ai[t]]++;
    ATp[0] = 0;
    for (int j = 0; j < n; j++)
    {
        ATp[j + 1] = ATp[j] + iwork[j];
        iwork[j] = ATp[j];
    }
    for (int i = 0; i < m; i++)
    {
        for (int t = Ap[i]; t < Ap[i + 1]; t++)
        {
            int dst = iwork[Ai[t]]++;
            ATi[dst] = i;
            if (Ax != NULL && ATx != NULL) ATx[dst] = Ax[t];
        }
    }
}

/* sparse_matrix.c:136-227: invert group by counting sort (sources of each
   output row ascending), output row = sorted union of source columns,
   map[source nnz] = output position. */
int orc_row_reduce_alloc(int m, int n, const int *Ap, const int *Ai,
                         const int *group, int m_out, int *Cp, int *Ci, int *map)
{
    int *start = (int *) calloc((size_t) m_out + 1, sizeof(int));
    int *cursor = (int *) malloc((size_t) (m_out > 0 ? m_out : 1) * sizeof(int));
    int *src = (int *) malloc((size_t) (m > 0 ? m : 1) * sizeof(int));
    int *where = (int *) malloc((size_t) (n > 0 ? n : 1) * sizeof(int));
    for (int i = 0; i < m; i++) start[group[i] + 1]++;
    for (int g = 0; g < m_out; g++) start[g + 1] += start[g];
    memcpy(cursor, start, (size_t) m_out * sizeof(int));
    for (int i = 0; i < m; i++) src[cursor[group[i]]++] = i;
    for (int j = 0; j < n; j++) where[j] = -1;

    int nnz = 0;
    Cp[0] = 0;
    for (int g = 0; g < m_out; g++)
    {
        int row_begin = nnz;
        for (int s = start[g]; s < start[g + 1]; s++)
        {
            int i = src[s];
            for (int t = Ap[i]; t < Ap[i + 1]; t++)
            {
                if (where[Ai[t]] < row_begin)
                {
                    where[Ai[t]] = nnz;
                    Ci[nnz++] = Ai[t];
                }
            }
        }
        qsort(Ci + row_begin, (size_t) (nnz - row_begin), sizeof(int), cmp_int);
        Cp[g + 1] = nnz;
        for (int e = row_begin; e < nnz; e++) where[Ci[e]] = e;
        for (int s = start[g]; s < start[g + 1]; s++)
        {
            int i = src[s];
            for (int t = Ap[i]; t < Ap[i + 1]; t++) map[t] = where[Ai[t]];
        }
    }
    free(start);
    free(cursor);
    free(src);
    free(where);
    return nnz;
}

/* sparse_matrix.c:229-240 with CSR_sum.c:150-157: zero, then out[map[j]] +=
   vals[j] for j ascending. */
void orc_row_reduce_fill(int nnzA, const double *Ax, const int *map, int nnzC,
                         double *Cx)
{
    memset(Cx, 0, (size_t) nnzC * sizeof(double));
    for (int t = 0; t < nnzA; t++) Cx[map[t]] += Ax[t];
}

/* row_reduce (sparse_matrix.c:136-240) applied to the triplets as one-entry
   rows grouped by their row index: stable counting sort by row, per output
   row the sorted unique union of columns, sums in triplet order. */
int64_t orc_coo_coalesce(int64_t nnz, const int *rows, const int *cols,
                         const double *x, int m, int n, int64_t *Cp, int *Ci,
                         double *Cx)
{
    int64_t *start = (int64_t *) calloc((size_t) m + 1, sizeof(int64_t));
    int64_t *order = (int64_t *) malloc((size_t) (nnz > 0 ? nnz : 1) * sizeof(int64_t));
    int64_t *where = (int64_t *) malloc((size_t) (n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t t = 0; t < nnz; t++) start[rows[t] + 1]++;
    for (int r = 0; r < m; r++) start[r + 1] += start[r];
    {
        int64_t *cursor = (int64_t *) malloc((size_t) (m > 0 ? m : 1) * sizeof(int64_t));
        memcpy(cursor, start, (size_t) m * sizeof(int64_t));
        for (int64_t t = 0; t < nnz; t++) order[cursor[rows[t]]++] = t;
        free(cursor);
    }
    for (int j = 0; j < n; j++) where[j] = -1;
    int64_t out = 0;
    Cp[0] = 0;
    for (int r = 0; r < m; r++)
    {
        int64_t row_begin = out;
        for (int64_t s = start[r]; s < start[r + 1]; s++)
        {
            int c = cols[order[s]];
            if (where[c] < row_begin)
            {
                where[c] = out;
                Ci[out++] = c;
            }
        }
        qsort(Ci + row_begin, (size_t) (out - row_begin), sizeof(int), cmp_int);
        for (int64_t e = row_begin; e < out; e++)
        {
            where[Ci[e]] = e;
            Cx[e] = 0.0;
        }
        for (int64_t s = start[r]; s < start[r + 1]; s++)
        {
            int64_t t = order[s];
            Cx[where[cols[t]]] += x[t];
        }
        Cp[r + 1] = out;
    }
    free(start);
    free(order);
    free(where);
    return out;
}

/* sparse_matrix.c:79-134: concatenate the selected rows (repeats allowed). */
int64_t orc_row_gather(int m_out, const int *map, const int *Ap, const int *Ai,
                       const double *Ax, int *Cp, int *Ci, double *Cx)
{
    Cp[0] = 0;
    for (int r = 0; r < m_out; r++)
    {
        int src = map[r];
        int len = Ap[src + 1] - Ap[src];
        memcpy(Ci + Cp[r], Ai + Ap[src], (size_t) len * sizeof(int));
        if (Ax != NULL && Cx != NULL)
            memcpy(Cx + Cp[r], Ax + Ap[src], (size_t) len * sizeof(double));
        Cp[r + 1] = Cp[r] + len;
    }
    return Cp[m_out];
}

/* CSR_matrix.c:86-97: C.x = A.x, then C.x[j] *= d[row]. */
void orc_da(int m, const int *Ap, const double *Ax, const double *d, double *Cx)
{
    for (int r = 0; r < m; r++)
        for (int t = Ap[r]; t < Ap[r + 1]; t++) Cx[t] = Ax[t] * d[r];
}

/* linalg_sparse_matmuls.c:25-50: y_b = A x_b, each row a left-to-right dot. */
void orc_kron_spmv(int m, int n, const int *Ap, const int *Ai, const double *Ax,
                   const double *x, double *y, int p)
{
    for (int b = 0; b < p; b++)
    {
        for (int i = 0; i < m; i++)
        {
            double s = 0.0;
            for (int t = Ap[i]; t < Ap[i + 1]; t++) s += Ax[t] * x[b * n + Ai[t]];
            y[b * m + i] = s;
        }
    }
}

/* COO_matrix.c:48-90: entries with col <= row, in CSR order. */
int orc_coo_lower(int m, const int *Ap, const int *Ai, const double *Ax, int *rows,
                  int *cols, double *x, int *value_map)
{
    int cnt = 0;
    for (int r = 0; r < m; r++)
    {
        for (int t = Ap[r]; t < Ap[r + 1]; t++)
        {
            if (Ai[t] <= r)
            {
                rows[cnt] = r;
                cols[cnt] = Ai[t];
                x[cnt] = Ax[t];
                value_map[cnt] = t;
                cnt++;
            }
        }
    }
    return cnt;
}

/* Matrix sums (CSR_sum.c:230-297 sum_n_csr_alloc, matrix_sum.c:112-180).
   Output row r holds the ascending union of the operands' columns in row r;
   map[off[k] + t] is the output position of operand k's entry t. */
int orc_sum_n_alloc(int m, int n, int nops, const int *P, const int *I,
                    const int *off, int *Cp, int *Ci, int *map)
{
    int *where = (int *) malloc((size_t) (n > 0 ? n : 1) * sizeof(int));
    for (int j = 0; j < n; j++) where[j] = -1;
    int nnz = 0;
    Cp[0] = 0;
    for (int r = 0; r < m; r++)
    {
        int row_begin = nnz;
        for (int k = 0; k < nops; k++)
        {
            const int *p = P + (size_t) k * (m + 1);
            const int *idx = I + off[k];
            for (int t = p[r]; t < p[r + 1]; t++)
            {
                if (where[idx[t]] < row_begin)
                {
                    where[idx[t]] = nnz;
                    Ci[nnz++] = idx[t];
                }
            }
        }
        qsort(Ci + row_begin, (size_t) (nnz - row_begin), sizeof(int), cmp_int);
        for (int e = row_begin; e < nnz; e++) where[Ci[e]] = e;
        for (int k = 0; k < nops; k++)
        {
            const int *p = P + (size_t) k * (m + 1);
            for (int t = p[r]; t < p[r + 1]; t++) map[off[k] + t] = where[I[off[k] + t]];
        }
        Cp[r + 1] = nnz;
    }
    free(where);
    return nnz;
}

/* Cx = 0, then Cx[map[t]] += X[t] over the concatenated operand values in
   operand order (matrix_sum.c:142-155 with CSR_sum.c:150-157).  With d non-NULL
   (nops x m row scales) each term is the rounded product d[k*m + r] * X[t]
   (CSR_sum.c:116-148: val += d1[row] * a; val += d2[row] * b). */
void orc_sum_n_fill(int m, int nops, const int *P, const int *off, const double *X,
                    const double *d, const int *map, int nnzC, double *Cx)
{
    memset(Cx, 0, (size_t) nnzC * sizeof(double));
    for (int k = 0; k < nops; k++)
    {
        const int *p = P + (size_t) k * (m + 1);
        for (int r = 0; r < m; r++)
        {
            for (int t = p[r]; t < p[r + 1]; t++)
            {
                double v = X[off[k] + t];
                if (d) v = d[(size_t) k * m + r] * v;
                Cx[map[off[k] + t]] += v;
            }
        }
    }
}

/* mini_numpy.c:74-89: v[row + j*m] = sum over col ascending of
   Y[j + col*k] * w[row + col*m], from 0.0. */
void orc_y_kron_i_vec(int m, int k, int n, const double *Y, const double *w, double *v)
{
    for (int j = 0; j < k; j++)
        for (int row = 0; row < m; row++)
        {
            double s = 0.0;
            for (int col = 0; col < n; col++) s += Y[j + (size_t) col * k] * w[row + (size_t) col * m];
            v[row + (size_t) j * m] = s;
        }
}

/* mini_numpy.c:91-105: v[j + col*k] = sum over row ascending of
   X[row + j*m] * w[row + col*m], from 0.0. */
void orc_i_kron_xt_vec(int m, int k, int n, const double *X, const double *w, double *v)
{
    for (int col = 0; col < n; col++)
        for (int j = 0; j < k; j++)
        {
            double s = 0.0;
            for (int row = 0; row < m; row++) s += X[row + (size_t) j * m] * w[row + (size_t) col * m];
            v[j + (size_t) col * k] = s;
        }
}
