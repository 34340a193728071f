/* sparsediff_linalg drop-in: vertical stacking.  Same signatures as the
 * reference's include/sparsediff_linalg/matrix_vstack.h:27-34. */
#ifndef SDL_AMD_MATRIX_VSTACK_H
#define SDL_AMD_MATRIX_VSTACK_H

#include "sparsediff_linalg/matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C = [ops[0]; ...; ops[n-1]] as a sparse matrix (any operand kinds). */
matrix *vstack_matrices_alloc(matrix *const *ops, int n);
/* Copies only operands whose values_version moved since the previous fill;
   bumps C's version iff it copied something. */
void vstack_matrices_fill_values(matrix *const *ops, int n, matrix *C);

#ifdef __cplusplus
}
#endif

#endif
