// Shared host-side pieces of the drop-in C ABI: host array helpers, the
// registry that attaches device state to raw structs created by an alloc,
// and reference-style fatal errors.
#pragma once

#include "sdl_internal.h"
#include "sparse_ops.h"
#include "spgemm.h"

#include <memory>

extern "C" {
#include "sparsediff_linalg/COO_matrix.h"
#include "sparsediff_linalg/CSC_matrix.h"
#include "sparsediff_linalg/CSR_matrix.h"
}

namespace sdl {

// Reference-style hard error: message on stderr, exit(1)
// (sparse_matrix.c:92-97, permuted_dense.c:258-263).
[[noreturn]] void api_exit(const char *fn, const char *msg);

// Checks a computed nnz fits the int ABI.
int checked_nnz(const char *fn, int64_t nnz);

// Device state bound to a raw struct by the alloc that produced it, looked
// up by the matching fill.  Keyed by the struct's address.
struct Attached
{
    virtual ~Attached() = default;
};
void attach(const void *owner, std::unique_ptr<Attached> st);
Attached *attached(const void *owner);
void detach(const void *owner);

template <typename T> T *attached_as(const void *owner, const char *fn)
{
    Attached *a = attached(owner);
    T *t = a ? dynamic_cast<T *>(a) : nullptr;
    if (!t) api_exit(fn, "output was not produced by the matching alloc");
    return t;
}

// Host array allocation used for ABI structs (freed by the free_* functions),
// counted in the tracked_alloc.h counters (g_allocated_bytes / g_peak_bytes).
int *host_ints(int64_t n, bool zero);
double *host_doubles(int64_t n);
void *host_malloc(size_t bytes);
void host_free(void *p);


} // namespace sdl
