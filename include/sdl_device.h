/* Device-resident extension of the sparsediff_linalg drop-in (no reference
 * counterpart).  The reference's fills read and write host arrays; callers
 * that keep operands in HBM (a Newton solver on the GPU, the batched refill
 * benchmark) use these entry points instead.  All pointers named d* are device
 * pointers; stream arguments are hipStream_t (NULL = the library stream).
 * Functions return 0 on success and nonzero on error (message in
 * sdl_last_error()).  Refills allocate nothing. */
#ifndef SDL_DEVICE_H
#define SDL_DEVICE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ---- */
void sdl_init(void);
const char *sdl_last_error(void);
void *sdl_library_stream(void);
void sdl_synchronize(void);
long long sdl_device_alloc_count(void);
long long sdl_device_free_count(void);
/* Device memory and timing helpers for callers without a HIP toolchain
   (events are recorded on the given stream, NULL = library stream). */
#include <stddef.h>
void *sdl_dev_alloc(size_t bytes);
void sdl_dev_free(void *p);
int sdl_memcpy_htod(void *dst, const void *src, size_t bytes);
int sdl_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int sdl_memset_d(void *dst, int value, size_t bytes);
void *sdl_event_create(void);
void sdl_event_record(void *ev, void *stream);
float sdl_event_elapsed_ms(void *start, void *stop);
void sdl_event_destroy(void *ev);
int sdl_device_name(char *buf, int len);
/* Streaming device copy kernel (16-byte aligned), the HBM-counter calibration
   reference: exactly `bytes` read and `bytes` written. */
int sdl_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);
/* Streaming-write calibration: `bytes` of constant stores (mode 0: 16 bytes
   per lane, non-temporal; 1: 16 bytes, ordinary; 2: 8 bytes, non-temporal). */
int sdl_write_kernel(void *dst, size_t bytes, int mode, void *stream);
/* A batched refill's store pattern alone: K rows of n doubles at stride ldc,
   thread e of a 256-entry tile storing column e of `K / groups` consecutive
   rows (groups = 1: every row), one 8-byte non-temporal store per row. */
int sdl_write_rows_kernel(double *dC, long long ldc, int K, long long n, int groups,
                          void *stream);
/* Sustained FP64 matrix-rate calibration: blocks workgroups of 4 wavefronts,
   each wavefront running nacc (1, 2, 4, 8, 16) independent accumulators
   through iters rounds of v_mfma_f64_16x16x4_f64 on registers, so
   blocks * 4 * iters * nacc * 2048 flops; out receives blocks * 256 doubles. */
int sdl_mfma_kernel(int nacc, int blocks, int iters, double *out, void *stream);

/* ---- C = B^T diag(d) A on device-resident CSR operands (the hot path of
   BTA_matrices_alloc / BTDA_matrices_fill_values for two sparse operands,
   reference src/matmul_dispatchers.c:335-375) ---- */
typedef struct sdl_btda_plan sdl_btda_plan;

enum
{
    SDL_SYM_EXACT = 0,     /* count pass + fill pass */
    SDL_SYM_PREDICTED = 1, /* count pass sized by the row-nnz predictor, exact fallback */
    /* or-ed into mode: build the pull schedule (per-entry term lists) whatever
       the terms / nnz ratio, as the batched, captured and range refills need
       it (the reference builds it iff terms <= 3 nnz; values are the same) */
    SDL_PLAN_PULL = 256,
    /* or-ed into mode: the by-rank rows (above the row-accumulator cut, 1024
       entries by default) find each term's entry by ranking the row's columns
       at every refill, instead of the plan recording every term's position
       (2 bytes per term; recorded when these rows' terms are at most 1.25x
       their entries).  Values are the same. */
    SDL_PLAN_RANK_REFILL = 512
};

/* A: m x nA CSR (dAp[m+1], dAi[nnzA]); B: m x nB CSR.  The plan copies the
   structure it needs; the caller's arrays may be freed afterwards. */
int sdl_btda_plan_create(int m, int nA, int nB, const int *dAp, const int *dAi,
                         const int *dBp, const int *dBi, int mode, sdl_btda_plan **out);
void sdl_btda_plan_destroy(sdl_btda_plan *plan);
long long sdl_btda_plan_nnz(const sdl_btda_plan *plan);
long long sdl_btda_plan_terms(const sdl_btda_plan *plan);
int sdl_btda_plan_is_pull(const sdl_btda_plan *plan);
/* Device row pointer (int64, nB+1 entries) and column indices of C. */
const long long *sdl_btda_plan_rowptr(const sdl_btda_plan *plan);
const int *sdl_btda_plan_colind(const sdl_btda_plan *plan);
/* Copy C's structure to host int32 arrays (fails if nnz > INT_MAX). */
int sdl_btda_plan_structure(const sdl_btda_plan *plan, int *hp, int *hi);
/* Symbolic-pass statistics (11 values): rows, terms, nnz, the fill pass's
   hash / LDS-bitmap / global-bitmap rows, overflow and predicted rows, then
   the first count pass's hash / LDS-bitmap / global-bitmap rows (the count
   pass sizes each row from min(terms, column span), or from its prediction). */
int sdl_btda_plan_stats(const sdl_btda_plan *plan, long long *out11);
/* Window rows of a plan (4 values): rows counted (first count pass) and
   filled in span windows -- rows whose columns all lie in a span of at most
   8192 columns, marked in an LDS array over that span instead of a hash
   table or a column bitmap; then, for row-accumulator plans, rows of at most
   1024 entries refilled one wavefront per row, and rows of more entries
   (without recorded positions) refilled in 2048- / 4096- / 8192-column
   windows. */
int sdl_btda_plan_window_rows(const sdl_btda_plan *plan, long long *out4);
/* Span windows in the symbolic phase of plans created afterwards: 1
   (default) on, 0 off (hash tables and column bitmaps only); -1 only
   queries.  Both give the same structure.  Returns the previous mode, or -1
   for an unknown one. */
int sdl_symbolic_window(int mode);
/* Lanes per L entry of the span-window kernels: 4, 8, 16, 32 or 64, or 0
   (default) to pick them from the product's mean A-row length; -1 only
   queries.  Returns the previous setting, or -1 for an invalid one. */
int sdl_symbolic_window_lanes(int lanes);
/* Row-accumulator refills (plans without a pull schedule): 1 (default) rows
   of at most 1024 entries one wavefront per row, with a dense LDS accumulator
   over the row's span when it is at most 1024 columns, else over its entries
   with the columns searched in LDS; and longer rows spanning at most 8192
   columns, when the plan records no term positions, one wavefront per row
   over a 2048-, 4096- or 8192-column window.  0: the workgroup accumulator with a
   column search and the by-rank kernel.  Read when a plan is created, -1 only
   queries.  Both give the same values bit for bit.  Returns the previous
   mode, or -1 for an unknown one. */
int sdl_rowacc_window(int mode);
/* Term expansion of the symbolic phase's LDS column bitmaps: 1 (default)
   groups of lanes per L entry striding over its A row (as the span windows),
   0 the load-balanced expansion (a binary search per term); -1 only queries.
   Both give the same structure.  Returns the previous mode, or -1 for an
   unknown one. */
int sdl_symbolic_lane_groups(int mode);
/* Refill schedule (5 values): pull schedule (0/1), by-rank rows (row-
   accumulator plans), recorded positions for them (0/1), then with recorded
   positions the listed terms of entries of several terms and the rows left
   to the dense accumulator. */
int sdl_btda_plan_schedule(const sdl_btda_plan *plan, long long *out5);

/* Refills of one plan share the plan's workspace (operand pairs, transposed
   d, dense accumulators, the rank kernel's spill list): issue all refills of
   a plan on one stream, or serialise them; two refills of the same plan in
   flight on different streams race.  Different plans are independent.

   One refill: dCx[nnz] = B^T diag(dd) A with dAx/dBx the CSR values of A and
   B (dd may be NULL = identity). */
int sdl_btda_refill(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                    const double *dd, double *dCx, void *stream);
/* K refills sharing A and B: refill r reads dD + r*ldd (length m) and writes
   dC + r*ldc (length nnz).  One launch. */
int sdl_btda_refill_batched(const sdl_btda_plan *plan, const double *dAx,
                            const double *dBx, const double *dD, long long ldd, int K,
                            double *dC, long long ldc, void *stream);

/* K refills as sdl_btda_refill_batched, allocating nothing and needing no
   sdl_btda_plan_reserve: the refill's workspace (the term operand pairs and d
   transposed to k-major) goes to a fixed 32 MiB device array of the library
   (sdl_static_scratch_bytes()), so 2 * terms + m * (K rounded up to 16)
   doubles must fit it; larger calls return an error and launch nothing.  No
   counterpart in the reference: for callers that cannot allocate (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Every call on one GPU shares the array and leaves its workspace there, so
   calls that overlap on one GPU (different streams, or host threads) race;
   serialise them.  Results are bit-identical to sdl_btda_refill_batched. */
int sdl_btda_refill_batched_static(const sdl_btda_plan *plan, const double *dAx,
                                   const double *dBx, const double *dD, long long ldd, int K,
                                   double *dC, long long ldc, void *stream);
long long sdl_static_scratch_bytes(void);

/* Refill of the entries [e0, e1) of C only (plans with a pull schedule):
   dCx is the whole value array, entries outside the range are untouched.
   One launch: the per-call path for one instance of a block-diagonal plan. */
int sdl_btda_refill_range(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                          const double *dd, double *dCx, long long e0, long long e1,
                          void *stream);

/* Alloc-phase reservation of the batched-refill workspace for up to Kmax
   refills per call (d transposed to k-major + term operand pairs: 8*m*Kmax +
   16*terms bytes).  Batched refills with K <= Kmax then use the interleaved
   kernel; larger K fall back to direct gathers.  Allocates; refills do not. */
int sdl_btda_plan_reserve(sdl_btda_plan *plan, int Kmax);
/* As sdl_btda_plan_reserve, with the bound values (sdl_btda_bind_values) kept
   in tile layout instead of operand pairs: each term's A value (8 bytes per
   term) and each L entry's value (8 bytes per entry of B), so bound refills run
   the tile form of the interleaved refill (variant 7) reading them without an
   index; no pair array is allocated (8*m*Kmax + 8*terms + 8*nnz(B) bytes).
   Needs a pull schedule (sdl_btda_plan_is_pull) over a single-block L; returns
   1 otherwise.  Same values bit for bit. */
int sdl_btda_plan_reserve_tile(sdl_btda_plan *plan, int Kmax);

/* Gather the term operand pairs of the current A / B values into the
   reserved workspace (call again whenever A or B values change). */
int sdl_btda_bind_values(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                         void *stream);
enum
{
    SDL_BATCH_VALUES_BOUND = 1, /* operand pairs already bound: skip the gather */
    SDL_BATCH_D_KMAJOR = 2      /* dD is k-major: refill r of row k at dD[k*ldd + r], with
                                   ldd >= K rounded up to 16 (rows are read in runs of 16) */
};
/* Batched refill with options; ev_start / ev_stop (hipEvent_t, may be NULL)
   are recorded on the stream right before / after the main refill kernel. */
int sdl_btda_refill_batched_ex(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                               const double *dD, long long ldd, int K, double *dC,
                               long long ldc, unsigned flags, void *ev_start, void *ev_stop,
                               void *stream);

/* Launch configuration of the batched refill: variant 0 = one thread per
   entry over all K; 1 / 2 = grid over (entry tile, chunk of rchunk refills)
   with 1 / 2 entries per thread; 3 = interleaved (needs the reservation),
   4 / 5 = interleaved with 2 / 4 entries per thread, 6 = interleaved reading
   the operand values through the term lists (unbound values only; bound
   calls use 3), 7 = interleaved tile form: per 256-entry tile the scale
   factors L value * d[k] of the tile rows' L entries for the rchunk refills
   formed once in LDS, then one L index and one A value per term (unbound
   values, a single-block L; else 3). */
typedef struct sdl_refill_cfg
{
    int variant;
    int rchunk;
} sdl_refill_cfg;
int sdl_btda_refill_batched_cfg(const sdl_btda_plan *plan, const double *dAx,
                                const double *dBx, const double *dD, long long ldd, int K,
                                double *dC, long long ldc, const sdl_refill_cfg *cfg,
                                void *stream);
/* Candidate configurations (returns the count, writes up to max). */
int sdl_refill_candidates(sdl_refill_cfg *out, int max);
/* Autotuner: times every candidate on this plan and these buffers (reps timed
   launches each, after 2 warm-ups), keeps the fastest in the plan for later
   sdl_btda_refill_batched calls, records it in the tuning table under the
   plan's shape signature (below), and reports it. */
int sdl_btda_autotune(sdl_btda_plan *plan, const double *dAx, const double *dBx,
                      const double *dD, long long ldd, int K, double *dC, long long ldc,
                      int reps, sdl_refill_cfg *best, float *best_us);
/* Pin a configuration (NULL: back to the default heuristic). */
int sdl_btda_set_refill_cfg(sdl_btda_plan *plan, const sdl_refill_cfg *cfg);
/* Kernel of the plan's single refills (sdl_btda_refill, one d; pull plans):
   1 = tile refill (each 256-entry tile forms the scale factor L value * d[k]
   of its rows' L entries once, then reads one L index and one A value per
   term; the default when available: a pull schedule over a single-block L),
   0 = per-term gathers (three indices and three gathers per term), -1 = query.
   Same values bit for bit.  Returns the previous kind (0 / 1), or -1 when
   kind 1 is not available or kind is unknown. */
int sdl_btda_plan_single_refill(sdl_btda_plan *plan, int kind);

/* ---- Tuning table (no counterpart in the reference, whose symbolic and
   numeric phases are one fixed CPU algorithm, src/sparse/gustavson.c:20-150)
   Every plan creation looks up (this GPU's architecture, the product's shape
   signature) in a per-process table; an entry supplies the thresholds below
   and, when recorded, the batched refill's launch shape.  The signature is
   floor(log2) of the output rows, the columns and the mean terms per output
   row.  The thresholds only choose between kernels: structures and values
   are identical under every valid setting. */
typedef struct sdl_sym_tune
{
    int hash_max;   /* largest LDS hash table, slots: 1024, 2048, 4096 or 8192 */
    int dens;       /* a row of bound u takes the LDS bitmap when u*dens >= its
                       words (1 .. 64) */
    int rowacc_cut; /* longest row of the LDS row-accumulator refill (64 .. 1024) */
    int rank_cut;   /* longest row of the by-rank refill (rowacc_cut < . <= 65536) */
} sdl_sym_tune;
/* The built-in thresholds (8192, 8, 1024, 65536). */
int sdl_sym_tune_default(sdl_sym_tune *out);
/* Plan creation with forced thresholds (NULL: as sdl_btda_plan_create, i.e.
   the table's entry or the built-in values); out-of-range values are an error. */
int sdl_btda_plan_create_tuned(int m, int nA, int nB, const int *dAp, const int *dAi,
                               const int *dBp, const int *dBi, int mode,
                               const sdl_sym_tune *tune, sdl_btda_plan **out);
/* The thresholds a plan was created with. */
int sdl_btda_plan_tune(const sdl_btda_plan *plan, sdl_sym_tune *out);
/* Tunes the thresholds on this product.  A setting's cost is plan creation
   plus, with dAx and dBx given (both or neither), one refill; the least of
   reps (allocation stalls only add time).
   Stage 1 sweeps hash_max x dens, stage 2 (row-accumulator plans, values
   given) rowacc_cut x rank_cut from stage 1's winner; a candidate replaces
   the incumbent (first the built-in values) only when at least 3 % cheaper.
   Records the winner in the table under the product's signature (keeping a
   recorded refill shape), reports it and the two costs, in ms.  Fails if a
   candidate's output structure differs. */
int sdl_btda_tune(int m, int nA, int nB, const int *dAp, const int *dAi, const int *dBp,
                  const int *dBi, int mode, const double *dAx, const double *dBx, int reps,
                  sdl_sym_tune *best, float *best_ms, float *default_ms);
/* Records a batched refill launch shape (one of sdl_refill_candidates) for
   the plan's signature. */
int sdl_btda_tune_record_refill(const sdl_btda_plan *plan, const sdl_refill_cfg *cfg);
/* Entries in the table; drop them all. */
int sdl_tune_table_size(void);
void sdl_tune_table_clear(void);
/* Save the table as text closed by an FNV-1a-64 checksum line; load merges a
   saved table over the current one (the file's entries win).  A file with a
   wrong header or checksum, or any out-of-range entry, is rejected whole and
   leaves the table unchanged.  Plans created before a load keep their
   thresholds. */
int sdl_tune_table_save(const char *path);
int sdl_tune_table_load(const char *path);

/* ---- COO coalescing on device-resident triplets ----
   The reference has no direct COO -> CSR entry; its coalescing semantics are
   row_reduce's (src/sparse/sparse_matrix.c:136-240): output rows hold sorted,
   unique columns and every entry is ((+0.0 + x_t1) + x_t2) + ... over its
   duplicates in input order.  drows / dcols: n_items int32 coordinates in an
   m x n space (n_items < 2^32). */
typedef struct sdl_coo_plan sdl_coo_plan;
int sdl_coo_plan_create(long long n_items, const int *drows, const int *dcols, int m, int n,
                        sdl_coo_plan **out);
long long sdl_coo_plan_nnz(const sdl_coo_plan *plan);
/* Device row pointer (int64, m+1) and column indices (nnz) of the output. */
const long long *sdl_coo_plan_rowptr(const sdl_coo_plan *plan);
const int *sdl_coo_plan_colind(const sdl_coo_plan *plan);
/* dCx[nnz] from the n_items values dx; one pass, allocates nothing. */
int sdl_coo_fill(const sdl_coo_plan *plan, const double *dx, double *dCx, void *stream);
void sdl_coo_plan_destroy(sdl_coo_plan *plan);

/* ---- Sums of same-shape sparse matrices on the device ----
   Device form of the reference's sum_n_matrices_alloc / _fill_values /
   _fill_add (src/sparse/matrix_sum.c:112-180, CSR_sum.c:230-297): the output
   holds the sorted union of the operands' columns per row, and every entry is
   ((+0.0 + x_0) + x_1) + ... over the operands holding it, in operand order
   (each operand's repeated entries in storage order).  dP[k] / dI[k]: device
   row pointers (m+1) and columns (nnz[k]) of operand k; the pointer arrays are
   host arrays.  keep_rows keeps each entry's row (4 bytes per operand entry)
   for the row-scaled fill. */
typedef struct sdl_sum_plan sdl_sum_plan;
int sdl_sum_plan_create(int nops, int m, int n, const int *const *dP, const int *const *dI,
                        const long long *nnz, int keep_rows, sdl_sum_plan **out);
long long sdl_sum_plan_nnz(const sdl_sum_plan *plan);
/* Device row pointer (int64, m+1) and column indices (nnz) of the output. */
const long long *sdl_sum_plan_rowptr(const sdl_sum_plan *plan);
const int *sdl_sum_plan_colind(const sdl_sum_plan *plan);
/* Device int32 map: output position of every operand entry, operands back to
   back (the reference's bound index maps, matrix_sum.c:90-110). */
const int *sdl_sum_plan_map(const sdl_sum_plan *plan);
/* 1 when the unscaled fill streams the operand values through per-entry operand
   masks (at most 16 operands, each with sorted unique columns per row); 0 when
   it reads them through the entry -> operand-entry lists. */
int sdl_sum_plan_is_masked(const sdl_sum_plan *plan);
/* The plan's fill: 0 = entry -> operand-entry lists, 1 = operand masks,
   2 = scatter (operands that rarely share an output entry: at most 32 operands,
   operand entries <= ratio * output entries, sdl_sum_scatter_ratio, and on
   average 16 or more consecutive output entries from one operand or at most
   4,194,304 output entries; each
   single-source entry written as 0.0 + x straight from its operand, in operand
   order, the shared entries summed through the lists); -1 for NULL. */
int sdl_sum_plan_fill_kind(const sdl_sum_plan *plan);
/* Ratio of operand entries to output entries up to which later sum plans use
   the scatter fill (default 1.25; 0 = never); ratio < 0 only queries.
   Returns the previous ratio. */
double sdl_sum_scatter_ratio(double ratio);
/* dCx[nnz] = the sum; dX: host array of nops device value pointers.  dD (device,
   nops x m, or NULL): row scales, each term the rounded product dD[k*m + row] * x
   (sum_scaled_matrices_fill_values, CSR_sum.c:116-148, for any count).  Up to 32
   operands the launch allocates and uploads nothing (graph-capturable). */
int sdl_sum_fill(sdl_sum_plan *plan, const double *const *dX, const double *dD, double *dCx,
                 void *stream);
/* dCx += operand k (device values dx): the incremental fill_add. */
int sdl_sum_add(sdl_sum_plan *plan, int k, const double *dx, double *dCx, void *stream);
void sdl_sum_plan_destroy(sdl_sum_plan *plan);

/* ---- Solver hand-off: lower-triangular COO of an HBM-resident output ----
   Device form of new_COO_matrix_lower_triangular / refresh_lower_triangular_coo
   (COO_matrix.c:48-98): the entries with col <= row in CSR order, value_map =
   their CSR positions (int64).  dCp (int64, m+1) and dCi describe a CSR with
   sorted columns per row, e.g. sdl_btda_plan_rowptr / _colind of a product
   plan; the refresh copies the selected values of dCx into dx (nnz of the
   export) without reading value_map, allocates nothing and can be captured. */
typedef struct sdl_lower_coo sdl_lower_coo;
int sdl_lower_coo_create(int m, const long long *dCp, const int *dCi, sdl_lower_coo **out);
long long sdl_lower_coo_nnz(const sdl_lower_coo *coo);
const int *sdl_lower_coo_rows(const sdl_lower_coo *coo);
const int *sdl_lower_coo_cols(const sdl_lower_coo *coo);
const long long *sdl_lower_coo_value_map(const sdl_lower_coo *coo);
int sdl_lower_coo_refresh(const sdl_lower_coo *coo, const double *dCx, double *dx, void *stream);
void sdl_lower_coo_destroy(sdl_lower_coo *coo);

/* ---- Hessian assembly across expression nodes (no reference counterpart) ----
   The caller-side loop of a Newton solver's Lagrangian Hessian, one
   evaluation per step: node k contributes C_k = A_k^T diag(d_k) A_k (the
   chain-rule product of BTDA_matrices_fill_values, matmul_dispatchers.c:73-92,
   for A_k: m[k] x n CSR, device arrays dAp[k] / dAi[k]); H = C_0 + ... +
   C_{nnodes-1} with sum_n semantics (matrix_sum.c:112-180: the union
   structure, each entry summed from +0.0 in node order); and H's lower
   triangle in the solver's COO layout (refresh_lower_triangular_coo,
   COO_matrix.c:92-98).  1 to 32 nodes.  The structures are built once at
   create; each evaluation runs every node's refill, the sum fill and the
   export as device launches, eagerly (sdl_hessian_eval) or replayed from one
   HIP graph captured over fixed buffers (sdl_hessian_capture, then
   sdl_hessian_replay per step: the caller rewrites the d_k / A_k value
   buffers in place between replays).  Neither allocates.  dAx[k]: A_k's
   values (device), dd[k]: m[k] doubles (device) or dd NULL (identity); dHx:
   H's values (sdl_hessian_nnz), dLx: the lower export's values
   (sdl_hessian_lower_nnz; may be NULL to skip the export). */
typedef struct sdl_hessian sdl_hessian;
int sdl_hessian_create(int nnodes, int n, const int *m, const int *const *dAp,
                       const int *const *dAi, sdl_hessian **out);
void sdl_hessian_destroy(sdl_hessian *h);
/* How the node refills are issued: 0 (default) = one launch over every node's
   256-entry tiles (up to 16 nodes per launch; nodes without the tile layout,
   e.g. row-accumulator schedules, fall back to 1), 1 = one launch per node in
   order on the caller's stream, 2 to 4 = one launch per node forked over that
   many streams.  Same values bit for bit; a later capture records the current
   setting. */
int sdl_hessian_set_streams(sdl_hessian *h, int streams);
long long sdl_hessian_nnz(const sdl_hessian *h);
const long long *sdl_hessian_rowptr(const sdl_hessian *h);
const int *sdl_hessian_colind(const sdl_hessian *h);
long long sdl_hessian_lower_nnz(const sdl_hessian *h);
const int *sdl_hessian_lower_rows(const sdl_hessian *h);
const int *sdl_hessian_lower_cols(const sdl_hessian *h);
int sdl_hessian_eval(sdl_hessian *h, const double *const *dAx, const double *const *dd,
                     double *dHx, double *dLx, void *stream);
int sdl_hessian_capture(sdl_hessian *h, const double *const *dAx, const double *const *dd,
                        double *dHx, double *dLx, void *stream);
int sdl_hessian_replay(const sdl_hessian *h, void *stream);

/* ---- Dense Kronecker-vector products on device arrays ----
   Device forms of the reference's Y_kron_I_vec / I_kron_XT_vec
   (src/support/mini_numpy.c:74-105): same index conventions, every output a
   left-to-right sum from +0.0 of individually rounded products.  No
   allocation; stream NULL = library stream. */
int sdl_y_kron_i_vec(int m, int k, int n, const double *dY, const double *dw, double *dv,
                     void *stream);
int sdl_i_kron_xt_vec(int m, int k, int n, const double *dX, const double *dw, double *dv,
                      void *stream);

/* ---- HIP graphs of library calls ----
   Work issued on `stream` (NULL = library stream) between sdl_graph_begin
   and sdl_graph_end is captured, not run; sdl_graph_launch replays it.  Only
   calls that allocate nothing and do not synchronize may be captured (the
   refills with their workspace reserved and operand pairs bound). */
typedef struct sdl_graph sdl_graph;
int sdl_graph_begin(void *stream);
int sdl_graph_end(void *stream, sdl_graph **out);
int sdl_graph_launch(const sdl_graph *graph, void *stream);
void sdl_graph_destroy(sdl_graph *graph);

/* ---- learned row binning (no reference counterpart) ----
   A 16 -> 32 -> 6 FP32 MLP predicts, from 16 structural features of an
   output row, the size bin of the row's nnz (<= 32, 128, 512, 2048, 4096,
   larger).  A predicted plan runs one symbolic pass with each row's hash
   table sized at twice its bin's bound (or from the term bound when that is
   smaller); rows whose table fills up take the exact count + fill passes, so
   the structure is exact whatever the prediction.  Parameters: W1[32][16] |
   b1[32] | W2[6][32] | b2[6] (742 floats).  dX: n x 16 features (row-major),
   dY: bins (int32). */
typedef struct sdl_mlp sdl_mlp;
int sdl_mlp_create(uint64_t seed, sdl_mlp **out);
void sdl_mlp_destroy(sdl_mlp *mlp);
int sdl_mlp_num_params(void);
int sdl_mlp_get_params(const sdl_mlp *mlp, float *params);
int sdl_mlp_set_params(sdl_mlp *mlp, const float *params);
/* Decision rule: the smallest bin whose cumulative probability >= tau
   (default 0.95). */
int sdl_mlp_set_tau(sdl_mlp *mlp, float tau);
/* Logits dZ[n][6]. */
int sdl_mlp_forward(const sdl_mlp *mlp, const float *dX, long long n, float *dZ);
int sdl_mlp_predict(const sdl_mlp *mlp, const float *dX, long long n, unsigned char *dbins);
/* Mean cross-entropy over the n rows and its gradient dG[742] (device). */
int sdl_mlp_gradients(sdl_mlp *mlp, const float *dX, const int *dY, long long n, float *dG,
                      float *loss);
/* Adam (lr) over shuffled mini-batches of `batch` rows for `epochs` passes;
   final_loss = mean batch loss of the last epoch. */
int sdl_mlp_train(sdl_mlp *mlp, const float *dX, const int *dY, long long n, int epochs,
                  int batch, float lr, uint64_t seed, float *final_loss);
/* Parameters to / from a file (magic, count, tau, parameters, FNV-1a
   checksum); load rejects a file whose checksum does not match. */
int sdl_mlp_save(const sdl_mlp *mlp, const char *path);
int sdl_mlp_load(sdl_mlp *mlp, const char *path);
/* Predictor for the sparse product allocs of the reference API and for
   sdl_btda_plan_create(..., SDL_SYM_PREDICTED, ...) (NULL: exact passes for
   the allocs; term-bound sizes for SDL_SYM_PREDICTED, i.e. the exact
   path).  Predicted symbolic phase: the count pass sizes each row's
   accumulator from the row's predicted nnz bin where the term bound would
   need a large column bitmap (at least 4x the predicted hash table, or one
   in global memory), rows whose count exceeds the bin are counted again from
   the term bound, and the fill pass sizes from the exact counts, so the
   structure is always exact. */
int sdl_set_row_predictor(const sdl_mlp *mlp);
/* Plan with the predicted symbolic phase (mlp NULL: term-bound sizes, the
   exact path). */
int sdl_btda_plan_create_predicted(int m, int nA, int nB, const int *dAp, const int *dAi,
                                   const int *dBp, const int *dBi, const sdl_mlp *mlp,
                                   sdl_btda_plan **out);
/* Features dF[nB][16] of the plan's output rows, and the bins of its exact
   row counts (training labels). */
int sdl_btda_plan_features(const sdl_btda_plan *plan, float *dF);
int sdl_btda_plan_row_bins(const sdl_btda_plan *plan, int *dY);
/* Predicted plans, over non-empty rows: counted within the predicted bound,
   counted again from the term bound (the bound was exceeded), predicted bin
   (tau rule) different from the bin of the exact nnz among the rows the
   network evaluated, bound not lowered (term bound kept), and the rows the
   network evaluated (those whose count a prediction could make at least 20 %
   cheaper; the others keep their term bound unexamined).  The bound of an
   evaluated row is lowered to the bin of least expected count cost, the
   chance of a second count weighed by the network's probabilities. */
int sdl_btda_plan_pred_stats(const sdl_btda_plan *plan, long long *out5);

/* ---- device refills of permuted_dense / stacked_pd product outputs ---- */
struct matrix;
/* Runs the kernels of a fill of C (an output of BTA_matrices_alloc,
   ATA_matrices_alloc or BA_matrices_alloc of block kind) with d already in
   device memory (dd: m doubles, or NULL for the d-free products): left is B
   for BTA / BA outputs and A for ATA outputs, right is A.  C's device values
   are written; its host values and values_version are left alone until
   sdl_block_publish.  Operand device mirrors are refreshed iff their version
   moved.  ev_start / ev_stop (hipEvent_t, may be NULL) bracket the kernels on
   the library stream.  Allocates nothing. */
int sdl_block_refill_device(struct matrix *C, struct matrix *left, struct matrix *right,
                            const double *dd, void *ev_start, void *ev_stop);
/* As sdl_block_refill_device, with ev_products (may be NULL) recorded after
   the block products (GEMM and sparse x dense lists) and before the ordered
   coalesce into C's blocks: the two phases timed apart. */
int sdl_block_refill_device_split(struct matrix *C, struct matrix *left, struct matrix *right,
                                  const double *dd, void *ev_start, void *ev_products,
                                  void *ev_stop);
/* out[8]: GEMM flops, sparse x dense flops, GEMM tiles, GEMM tasks, raw
   partial cells, output cells, coalesce terms, sparse x dense lists. */
int sdl_block_stats(const struct matrix *C, double *out);
/* Device pointer of C's values (block-major for stacked_pd). */
double *sdl_block_device_values(struct matrix *C);
/* Copies C's device values to its host values and bumps values_version. */
int sdl_block_publish(struct matrix *C);
/* Reduction slice depth of the block-product GEMM (16 or 32, default 16);
   returns the previous depth, or -1 for an unsupported value. */
int sdl_block_gemm_slice(int sk);
/* How the next block-product allocs coalesce per-block partials into a
   stacked_pd output (stacked_pd_coalesce.c:326-381 sums each cell over its
   source blocks in ascending order, from +0.0): 0 (default) = by output row
   where the output has at most 8192 columns and every partial's row and column
   lists are distinct (a dense LDS row per output row, no per-cell lists),
   else, or with 1, per-cell pull lists.  Both give the same values bit for bit.
   Returns the previous mode, or -1 for an unknown one. */
int sdl_block_coalesce_mode(int mode);

/* ---- synthetic workload generators (host; deterministic in seed) ---- */
/* Row lengths uniform in [lo, hi]; writes Ap[m+1]; returns nnz. */
long long sdl_gen_row_lengths_uniform(uint64_t seed, int m, int lo, int hi, int *Ap);
/* Columns: distinct uniform in [0, n), sorted per row; values N(0,1).  Ax may
   be NULL. */
void sdl_gen_fill_rows(uint64_t seed, int m, int n, const int *Ap, int *Ai, double *Ax);
/* Uniform doubles in [lo, hi). */
void sdl_gen_uniform(uint64_t seed, long long count, double lo, double hi, double *out);
/* Standard normal doubles (Box-Muller). */
void sdl_gen_normal(uint64_t seed, long long count, double *out);

#ifdef __cplusplus
}
#endif

#endif
