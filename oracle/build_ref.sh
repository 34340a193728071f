#!/bin/bash
# Builds the reference library's BLAS-free sparse subset from its own sources,
# where they lie under /root/reference, into oracle/_ref/libref_sparse.so.
# Used only as a checker for the C restatement in oracle/ and as the
# "reference" CPU baseline.  The permuted_dense / stacked_pd sources need
# <cblas.h>, which this image lacks, so they are not built.
set -euo pipefail
REF=${SDL_REFERENCE_ROOT:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -d "$REF/src/sparse" ]; then
  echo "build_ref: $REF not present; keeping any prebuilt $OUT/libref_sparse.so"
  exit 0
fi
mkdir -p "$OUT"
SRCS="$REF/src/sparse/gustavson.c $REF/src/sparse/CSC_matrix.c $REF/src/sparse/CSR_matrix.c \
      $REF/src/sparse/COO_matrix.c $REF/src/sparse/CSR_sum.c $REF/src/sparse/sparse_matrix.c \
      $REF/src/sparse/linalg_sparse_matmuls.c $REF/src/support/utils.c $REF/src/support/tracked_alloc.c \
      $REF/src/support/mini_numpy.c"
gcc -std=c99 -O3 -DNDEBUG -fPIC -shared -Wl,-Bsymbolic -I"$REF/include" -I"$REF/src" $SRCS \
    -o "$OUT/libref_sparse.so.tmp"
mv "$OUT/libref_sparse.so.tmp" "$OUT/libref_sparse.so"
echo "build_ref: built $OUT/libref_sparse.so"
