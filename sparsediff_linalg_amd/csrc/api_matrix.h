// Private layouts of the matrix kinds behind the drop-in `matrix` handle
// (include/sparsediff_linalg/matrix.h).  Each kind embeds `matrix` first, keeps
// the ABI's host arrays as the truth callers read and write, and mirrors them
// on the device: structure once, values whenever values_version moved since
// the last upload.  State an alloc binds for its fill lives in `bound`.
#pragma once

#include "api_common.h"

#include <cstdint>
#include <memory>
#include <vector>

extern "C" {
#include "sparsediff_linalg/matmul.h"
#include "sparsediff_linalg/matmul_kron.h"
#include "sparsediff_linalg/matrix.h"
#include "sparsediff_linalg/matrix_vstack.h"
}

namespace sdl {

// Device-side state bound to an output by the alloc that produced it.
struct Bound
{
    virtual ~Bound() = default;
};

// Device mirror of one value buffer (owner's host `x` of `n` doubles).
struct ValMirror
{
    DBuf<double> x;
    uint64_t seen = 0;
    bool valid = false;
};

struct SparseM
{
    matrix base;                 // must stay first
    CSR_matrix *csr = nullptr;   // host truth, owned
    // device mirror
    bool dev = false;
    DBuf<int> dp, di;
    ValMirror val;
    DBuf<int> drow;              // row index of every entry (lazy, alloc phase)
    // CSC mirror structure: columns as rows, entry q at CSR position cperm[q]
    bool has_csc = false;
    DBuf<int> cp, ci, cperm;
    std::unique_ptr<Bound> bound;
    // CSR view of a permuted_dense / stacked_pd owner: device values come
    // from the owner (refreshed by ext_values), the view is freed with it.
    const double *(*ext_values)(SparseM *) = nullptr;
    void *ext_owner = nullptr;
};

SparseM *as_sparse(const matrix *A);
matrix *make_sparse(CSR_matrix *csr);

// Upload structure (once) and allocate the value mirror.
void sparse_ensure_dev(SparseM *S);
// Structure of the CSC mirror (transpose) on the device (alloc phase).
void sparse_ensure_csc(SparseM *S);
// Row index array (alloc phase).
void sparse_ensure_rows(SparseM *S);
// Values on the device, uploaded iff values_version moved (fill phase; the
// buffer exists from sparse_ensure_dev).
const double *sparse_dev_values(SparseM *S);
// After a fill wrote S->val.x on the device: copy to the host x, bump the
// version once and mark the mirror fresh.
void sparse_publish(SparseM *S);

[[noreturn]] void unsupported(const char *fn, const matrix *A, const matrix *B);

} // namespace sdl
