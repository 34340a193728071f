// Private layouts of the permuted_dense and stacked_pd kinds.
//
// permuted_dense (reference permuted_dense.h:29-78): a dense row-major m0 x n0
// block X at strictly increasing row_perm / col_perm of an m x n matrix.
// stacked_pd (reference stacked_pd.h:28-54): row-disjoint PD blocks whose
// values live in one block-major buffer owned by the stack.  Both keep the
// host values as the ABI truth and a device mirror uploaded iff
// values_version moved; blocks of a stack are views into its buffers.
// Structure (perms, signatures, unions) is host-side integer bookkeeping over
// the perms; every value is computed on the GPU.
#pragma once

#include "api_pd.h"
#include "pd_kernels.h"

#include <vector>

namespace sdl {

struct SPDM;

struct PDM
{
    matrix base;  // must stay first
    int m0 = 0, n0 = 0;
    std::vector<int> row_perm, col_perm, row_inv, col_inv;
    double *X = nullptr;     // host values (owned unless parent)
    SPDM *parent = nullptr;  // set when this PD is a block of a stack
    int64_t poff = 0;        // offset of X in the parent's buffers
    DBuf<double> dX;         // device values of a standalone PD
    uint64_t dseen = 0;
    bool dvalid = false;
    matrix *csrv = nullptr;  // CSR view (a sparse_matrix aliasing X)
    std::unique_ptr<Bound> bound;
    DBuf<int> drow;          // global row of every value (DA), lazy
};

struct SPDM
{
    matrix base;  // must stay first
    std::vector<PDM *> blocks;
    std::vector<int> src_p, src_idx;
    std::vector<int64_t> boff;
    double *x = nullptr;     // host block-major values (owned)
    DBuf<double> dx;
    uint64_t dseen = 0;
    bool dvalid = false;
    matrix *csrv = nullptr;  // CSR view
    PullList csr_gather;     // CSR position -> buffer position
    uint64_t csr_seen = 0;
    bool csr_valid = false;
    std::unique_ptr<Bound> bound;
    DBuf<int> drow;
};

PDM *as_pd(const matrix *A);
SPDM *as_spd(const matrix *A);

// Structure of a permuted_dense block (used for raw partials and outputs).
struct Blk
{
    int m = 0, n = 0;
    std::vector<int> rp, cp;
    int m0() const { return (int) rp.size(); }
    int n0() const { return (int) cp.size(); }
    long long size() const { return (long long) rp.size() * (long long) cp.size(); }
};

PDM *new_pdm(const Blk &b, const double *X);
SPDM *new_spdm(int m, int n, const std::vector<PDM *> &blocks, const std::vector<int> &src_p,
               const std::vector<int> &src_idx);

// Device values (fill phase; uploads iff the owner's version moved).
const double *pd_vals(PDM *A);
const double *spd_vals(SPDM *A);
const double *mat_vals(matrix *A);   // PD, SPD or sparse
// Output buffer and publication (d2h + version bump once).
double *out_buf(matrix *C);
void out_publish(matrix *C);

// Coalesce of raw partial blocks into row-disjoint output blocks
// (reference stacked_pd_coalesce.c:244-381): rows grouped by signature
// (ascending list of raw blocks touching the row), signatures catalogued in
// row-scan order, col_perm = sorted union.  Returns the output block
// structures, the src map, and per output cell the raw positions feeding it
// in raw-block order.
struct Coalesced
{
    std::vector<Blk> out;
    std::vector<int> src_p, src_idx;
    std::vector<int> seg, idx;   // pull lists over the output's block-major cells
};
// lists = false leaves seg / idx empty (the row coalesce needs neither).
Coalesced coalesce_blocks(int m, int n, const std::vector<Blk> &raw,
                          const std::vector<long long> &raw_off, bool lists = true);

// Whether the row coalesce applies to these raw blocks: n <= ROWC_MAX_N, and
// every raw block's row and column lists hold distinct in-range indices.
bool rowc_eligible(int m, int n, const std::vector<Blk> &raw);
// The row coalesce of raw blocks into co's output blocks (co from
// coalesce_blocks on the same raw blocks).
void rowc_build(RowCoalesce &P, int m, int n, const std::vector<Blk> &raw,
                const std::vector<long long> &raw_off, const Coalesced &co);

// helpers
bool has_overlap(const std::vector<int> &a, const std::vector<int> &b);
int sorted_intersect(const std::vector<int> &a, const std::vector<int> &b, std::vector<int> &ia,
                     std::vector<int> &ib);
std::vector<int> sorted_union(const std::vector<const std::vector<int> *> &arrs);

} // namespace sdl
