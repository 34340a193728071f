// Generic device kernels behind the permuted_dense / stacked_pd kinds.  Every
// PD/SPD operation is an alloc-time descriptor set run by one of them:
//   * pull_sum:  out[e] = ((0 + src[idx[s0]]) + src[idx[s0+1]]) + ... (ordered);
//   * sdmm:      sparse x dense with per-list ordered sums (pd x csc, csc x pd,
//                pd x csr families);
//   * ggemm:     grouped gathered GEMM on FP64 MFMA (every dense-block product).
#pragma once

#include "sdl_internal.h"

#include <vector>

namespace sdl {

// ---------------------------------------------------------------- pull ----
// out[e] = (0 + src[idx[seg[e]]]) + src[idx[seg[e] + 1]] + ... for n outputs.
// Outputs with exactly one source, in runs of consecutive outputs reading
// consecutive sources, are copied by chunks (out + co[c] <- 0 + src + cs[c],
// cl[c] cells); the rest are listed with their output positions.
struct PullList
{
    int n = 0;                   // outputs
    int nchunks = 0;
    DBuf<int> co, cs, cl;        // chunk output start, source start, length
    int nrest = 0;
    DBuf<int> pos, seg, idx;     // rest: pos[nrest] (absent when the rest is every
                                 // output in order), seg[nrest+1], idx (sources in order)
};
void pull_build(PullList &P, const std::vector<int> &seg, const std::vector<int> &idx);
void pull_sum(const PullList &P, const double *src, double *out, hipStream_t s);

// --------------------------------------------------------- row coalesce ----
// The ordered coalesce of raw partial blocks by output row, with no per-cell
// lists: output row r's sources are the rows of the raw blocks touching r, in
// ascending block order, each a contiguous run of the block's n0 cells whose
// global columns are the block's column list.  One workgroup per output row
// adds them source by source into a dense LDS row over the n global columns,
// from +0.0, marking the columns touched, and writes the touched columns in
// ascending order: the row of its signature block, whose columns are the
// sorted union of those lists.  Each cell is therefore the same left fold in
// ascending block order as the pull lists give.  Needs n <= ROWC_MAX_N and
// distinct row and column lists within every raw block.
constexpr int ROWC_MAX_N = 8192;
struct RowCoalesce
{
    int nrows = 0;              // output rows (block-major output order)
    int ncolw = 0;              // n rounded up to a multiple of 32
    int h = 1;                  // 256-cell chunks of the longest source row
    DBuf<long long> obase;      // output offset of each row
    DBuf<int> sp;               // nrows + 1: source range of each row
    DBuf<long long> sstart;     // raw offset of each source row
    DBuf<int> slen, scp;        // its cell count and column-list offset in cpool
    DBuf<int> cpool;            // every raw block's column list
    long long cells = 0;        // raw cells read per coalesce
};
void rowc_run(const RowCoalesce &P, const double *raw, double *out, hipStream_t s);
// 0: the row coalesce where it applies (default), 1: always the pull lists.
extern int g_coalesce_mode;

// ---------------------------------------------------------------- sdmm ----
// Output list L (a row or a column of the output block) sums its entries in
// order: entry t = (dense index dr, sparse value position vp, d index dk or -1).
// dense(c) = D[dr*ld + c] (or D[c*ld + dr] when dense_trans).
// assoc 0: term = (d[dk] * S[vp]) * dense(c);  assoc 1: term = S[vp] * (d[dk] * dense(c)).
// C[L*sL + c*sC] for c < NC.
struct SdmmPlan
{
    int NL = 0, NC = 0;
    int ld = 0;
    bool dense_trans = false;
    int assoc = 0;
    long long sL = 0, sC = 0;
    DBuf<int> lp;                 // NL + 1
    DBuf<int> dr, vp, dk;         // entries
    long long nent = 0;           // number of entries
    bool has_dk = false;
};
void sdmm_build(SdmmPlan &P, const std::vector<int> &lp, const std::vector<int> &dr,
                const std::vector<int> &vp, const std::vector<int> &dk);
void sdmm_run(const SdmmPlan &P, const double *Sx, const double *D, const double *d,
              double *C, hipStream_t s);

// ---------------------------------------------------------------- ggemm ----
// C[i][j] = sum over the job's tasks (in order) of sum_s P_t(s, i) * Q_t(s, j),
//   P_t(s,i) = p_mode 0: P[p_off + prow[s]*p_ld + pmap(i)]  |  1: P[p_off + pmap(i)*p_ld + prow[s]]
//   Q_t(s,j) = w(s) * (q_mode 0: Q[q_off + qrow[s]*q_ld + qmap(j)] | 1: Q[q_off + qmap(j)*q_ld + qrow[s]])
// with w(s) = d[widx[s]] (scaled before the product) when widx >= 0 (on the
// device widx is rewritten to a position in the plan's gathered w); map
// entries < 0 contribute zero; a map offset < 0 is the identity.  P and Q
// read source buffer 0 or 1 (p_buf / q_buf).
struct GTask
{
    long long p_off, q_off;
    int p_buf, q_buf;
    int p_ld, q_ld;
    int p_mode, q_mode;
    int S;
    int prow, pmap;
    int qrow, qmap;
    int widx;
    // set by gemm_upload: bit 0 (1) = P's rows are p_r0 + s and pmap is the
    // identity, so P loads need no pool lookup; same for Q
    int direct = 0;
    int p_r0 = 0, q_r0 = 0;
};

struct GJob
{
    long long c_off;
    int ldc, M, N;
    int i0, j0;
    int t0, t1;
};

struct GemmPlan
{
    int njobs = 0;
    long long flops = 0;   // 2 * sum over tasks of S * M * N (algorithmic)
    DBuf<GTask> tasks;
    DBuf<GJob> jobs;
    DBuf<int> pool;
    // the d factors of every task, gathered once per refill (w[q] = d[wsrc[q]])
    // so that the GEMM's operand loads do not wait on a d-index load
    int nw = 0;
    DBuf<int> wsrc;
    DBuf<double> w;
};

// Host-side builder: add output blocks and their tasks, then upload.
struct GemmBuilder
{
    std::vector<GTask> tasks;
    std::vector<GJob> jobs;
    std::vector<int> pool;
    long long flops = 0;
    int add_ints(const std::vector<int> &v);  // returns the offset
    // One output block (M x N at c_off, row stride ldc) fed by tasks [t0, t1).
    void add_output(long long c_off, int ldc, int M, int N, int t0, int t1);
};
void gemm_upload(GemmPlan &P, const GemmBuilder &B);
void gemm_run(const GemmPlan &P, const double *buf0, const double *buf1, const double *d,
              double *C, hipStream_t s);

} // namespace sdl
