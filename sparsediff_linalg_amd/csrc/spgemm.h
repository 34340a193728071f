// Device-resident sparse x sparse product C = kron(I_blocks, L) . diag(d) . A
// (the reference's Gustavson core, src/sparse/gustavson.c), split into an
// alloc-phase plan (exact structure + fill schedule) and refills.
#pragma once

#include "sdl_internal.h"

#include <vector>

namespace sdl {

// Rows of a compressed matrix on the device: row r has entries
// (Li[q], value) for Lp[r] <= q < Lp[r+1].  The value of entry q lives at
// Lval[Lperm ? Lperm[q] : q] of the value array the refill is given, so a
// CSC mirror needs no value copy: Lperm maps its entries to the owner's CSR.
struct LView
{
    int Lm = 0;
    const int *Lp = nullptr;
    const int *Li = nullptr;
    const int *Lperm = nullptr;
    int blocks = 1;
};

struct AView
{
    int Am = 0, An = 0;
    const int *Ap = nullptr;
    const int *Ai = nullptr;
};

enum class SymMode
{
    Exact = 0,      // count pass + fill pass, tables sized from the term bound / exact count
    Predicted = 1,  // count pass with tables sized from a per-row prediction; rows
                    // over their predicted bound are counted again from the term bound
};

// Span windows in the symbolic phase: 0 off, 1 on (default).
extern int g_sym_window;
// Lanes per L entry of the window kernels (0: from the mean A-row length).
extern int g_sym_window_lanes;
// Lane groups per L entry in the LDS column-bitmap kernels: 0 off (the
// load-balanced term expansion), 1 on (default).
extern int g_sym_groups;
// Window accumulators for row-accumulator rows of narrow column span: 0 off,
// 1 on (default); read when a plan is built.
extern int g_rowacc_window;

struct SymStats
{
    long long rows = 0, terms = 0, nnz = 0;
    long long hash_rows = 0, bitmap_rows = 0, global_rows = 0;
    long long overflow_rows = 0;       // predicted path: count exceeded the predicted bound
    long long predicted_rows = 0;      // predicted path: counted within the predicted bound
    long long misbin_rows = 0;         // predicted path: predicted bin != bin of the exact nnz
    long long exact_rows = 0;          // predicted path: term bound kept (not lowered)
    long long candidate_rows = 0;      // predicted path: rows given features and the network
    long long count_hash_rows = 0, count_bitmap_rows = 0, count_global_rows = 0;  // first count pass
    long long window_rows = 0, count_window_rows = 0;  // span windows: fill pass, first count pass
};

// Launch configuration of the batched pull refill (picked by the autotuner):
// variant 0 = one thread per entry over all K refills; 1 = r-chunked grid, one
// entry per thread; 2 = r-chunked grid, two entries per thread (16-byte stores);
// 3 = d transposed to k-major and term operands pre-gathered (needs the
// reserved workspace), r-chunked grid; 4 / 5 = 3 with 2 / 4 entries per thread;
// 6 = 3 with the operands read through tA / tB in the kernel (no pair array;
// calls with bound values use 3).
struct RefillCfg
{
    int variant = -1;
    int rchunk = 8;
    RefillCfg() = default;
    RefillCfg(int v, int r) : variant(v), rchunk(r) {}
};

// Tuning parameters of the symbolic phase and the row-accumulator refill:
// the largest LDS hash table (slots; 1024 .. 8192), the density factor of
// the hash / bitmap choice (a row of bound u takes the LDS column bitmap when
// u * dens >= the bitmap's words), the longest row (entries) refilled by the
// LDS row accumulator before the by-rank kernel takes over (<= 1024), and the
// longest row refilled by rank before the dense accumulator takes over
// (<= 65536).
struct SymTune
{
    int hash_max = 8192;
    int dens = 8;
    int rowacc_cut = 1024;
    int rank_cut = 65536;
};
// true when every field is in the range above (hash_max a power of two)
bool tune_valid(const SymTune &t);

// Shape signature of a product for the tuning table: floor(log2) of the
// output rows, the columns and the mean terms per output row.
struct ShapeSig
{
    int lg_rows = 0, lg_cols = 0, lg_tpr = 0;
};
// The tuning table (per process; saved to / loaded from a checksummed file
// by the sdl_tune_table_* entry points): thresholds and refill launch shape
// recorded for (GPU architecture, shape signature).  lookup returns false
// when there is no entry (the defaults apply).
bool tune_lookup(const ShapeSig &sig, SymTune *tune, RefillCfg *refill);
void tune_record(const ShapeSig &sig, const SymTune *tune, const RefillCfg *refill);
// Thresholds forced on the plans this thread creates (the tuner's candidates).
extern thread_local const SymTune *g_force_tune;

// One row of the rank bin as its position refill reads it (built by the plan).
struct RankRow
{
    int64_t cbase;   // first entry of the row in C
    int64_t tbase;   // first term of the row in the recorded positions
    int64_t mb;      // first item of the row in the item positions
    int q0, q1, koff, cnt, nl, skip;
};

struct ProductPlan
{
    LView L;
    AView A;
    int64_t m_out = 0, n = 0, nnz = 0, terms = 0;
    DBuf<int64_t> Cp;   // m_out + 1
    DBuf<int> Ci;       // nnz
    // Pull schedule (terms <= 3 * nnz, the reference's plan policy): the terms
    // of entry e are [ep[e], ep[e+1]) in ascending k.
    bool pull = false;
    DBuf<int64_t> ep;
    DBuf<int> tA, tB, tK;
    // Tile refill (single-block L): each term's L index q (tB = Lperm[q], tK =
    // Li[q]), and per 256-entry tile the L entries of its rows, so a single
    // refill forms each L entry's scale factor once per tile.  tile_refill:
    // single refills use it (else the per-term gathers).
    DBuf<int> tq;
    DBuf<int2> tqr;
    bool tile_refill = false;
    // Bound values in tile layout (a tile reservation instead of operand
    // pairs): ta[s] = A value of term s, lv[q] = L value of L entry q.
    DBuf<double> ta, lv;
    int64_t nl = 0;
    // Row-accumulator schedule (otherwise): rows binned by nnz -- [0, 1) up to
    // 1024 entries (LDS accumulators), [1, 2) up to 65536 (by rank when the
    // column bitmap fits LDS), [2, 3) longer.
    DBuf<int> rbins;                 // row ids grouped by class
    int rbin_off[8] = {0};           // class c is rbins[rbin_off[c] .. rbin_off[c+1])
    // bin 0 split for the wavefront accumulators: [0, rwin_off[1]) the
    // workgroup accumulator, then classes 2..4 the windows of 256 / 512 /
    // 1024 columns, classes 5..7 the column searches over 256 / 512 / 1024
    // entries (class c is rbins[rwin_off[c - 1] .. rwin_off[c]))
    int rwin_off[8] = {0};
    // rank-bin rows (no recorded positions) in the 2048- / 4096- / 8192-column
    // window accumulators: the bin's first rwide_off[3] rows, by window size
    int rwide_off[4] = {0};
    double mean_arow = 0;            // terms per L entry (the mean A-row length)
    bool rank = false;               // bin 1 by rank; its spills in spill[0 .. *nspill)
    int rank_grid = 0;
    DBuf<int> spill, nspill;
    // Rows too long for LDS accumulate in a dense per-workgroup accumulator of
    // n doubles (reserved here: dense_grid * n), or in C itself when n is too
    // large for that (dense_grid == 0).
    int dense_grid = 0;
    DBuf<double> acc;
    // Rank bin with recorded positions (its terms at most 1.25x its entries;
    // the plan records them unless rank_refill): every term's entry position
    // within its row, or the item index of a term of an entry of several
    // terms, whose items are the entry positions sorted by (entry, L index);
    // rows of more than 32767 entries or 1024 items are marked in rskip and
    // listed in spill[0 .. *nspill) once, by the plan.  A refill stages each
    // row in LDS windows of rank_window entries.
    bool rank_refill = false;        // set before planning: keep per-refill ranks
    bool rankpos = false;
    int rank_window = 0;
    DBuf<int64_t> roff;              // rank-bin row ri's terms start at roff[ri]
    DBuf<uint16_t> rpos;
    DBuf<unsigned char> rskip;
    int64_t nmulti = 0;              // items
    DBuf<int64_t> moff;              // row ri's items start at moff[ri] ...
    DBuf<int> mcnt;                  // ... and number mcnt[ri]
    DBuf<uint16_t> ipos;
    DBuf<RankRow> rrows;             // the refill's per-row descriptors
    DBuf<int2> qrange;               // one-block L: each L entry's A-row (start, length)
    SymStats stats;
    // Span windows of the symbolic phase (set while the plan is built, then
    // cleared): each output row's smallest column, column span and term
    // count, and the lanes per L entry of the window kernels.
    const int *win_lo = nullptr, *win_span = nullptr;
    const int64_t *win_terms = nullptr;
    int sym_sg = 64;
    RefillCfg tuned;   // variant < 0: not tuned, use the default heuristic
    SymTune tune;      // thresholds the symbolic phase and the refill bins used
    ShapeSig sig;
    // Batched-refill workspace (reserved in the alloc phase): term operand pairs
    // in term order, and D transposed to k-major with row stride batch_ld.
    DBuf<double> tv, dT;
    int batch_cap = 0;
    int64_t batch_ld = 0;
};

// Workspace of the interleaved batched refill given by the caller instead of
// the plan's reservation: operand pairs (2 * terms doubles) and d transposed
// to k-major (m * ld doubles, ld = K rounded up to 16).
struct BatchScratch
{
    double *tv = nullptr;
    double *dT = nullptr;
    int64_t ld = 0;
};
// Doubles of workspace the interleaved refill of K refills needs.
int64_t batch_scratch_doubles(const ProductPlan &P, int K);

// Reserve the interleaved batched-refill workspace for up to Kmax refills:
// the transposed d and, for bound values, operand pairs (16 B per term) or,
// with tile (single-block pull plans), the tile layout (8 B per term and per
// L entry).
void product_reserve_batch(ProductPlan &P, int Kmax, bool tile = false);

// Single refills of n plans in one launch per 16 plans (the tile refill's
// kernel over every plan's tiles).  Returns false, launching nothing, unless
// every plan is a pull plan refilled by tiles (sdl_btda_plan_single_refill).
// d may be null (no d for any plan) or hold a null per plan.
bool product_numeric_multi(const ProductPlan *const *plans, int n, const double *const *Lval,
                           const double *const *d, const double *const *Ax, double *const *Cx,
                           hipStream_t s);

// One refill of the entries [e0, e1) only (pull plans): Cx[e0..e1) written,
// the rest of Cx untouched.  One launch.
void product_numeric_range(const ProductPlan &P, const double *Lval, const double *d,
                           const double *Ax, double *Cx, int64_t e0, int64_t e1, hipStream_t s);

struct RowPredictor;

// Build the exact structure and the refill schedule (pull iff terms <= 3 nnz,
// the reference's plan policy; force_pull: whenever the counts fit int32 --
// both schedules give the same values bit for bit).  SymMode::Predicted: the
// count pass sizes each row's accumulator from its predicted bin where the term
// bound would need a large column bitmap (pred may be null: the term bound,
// i.e. the exact path), rows
// whose count exceeds the predicted bound are counted again from the term
// bound, and the fill pass is the exact path's.
void product_plan(ProductPlan &P, const LView &L, const AView &A, SymMode mode,
                  const RowPredictor *pred, bool allow_pull = true, bool force_pull = false);

// The 16 structural features of every output row (F[row * 16 + f], FP32):
// log2(1 + x) of the term bound, the degree, the longest / shortest A row,
// the mean A row, the column span of the row's A rows, the uniform-collision
// estimate within that span, span (1 - e^(-terms/span)), the summed
// single-row column spans, min(terms, summed spans), min(terms, span), the
// A-row index span, degree / index span, the root of the sum of squared
// A-row lengths, the mean single-row column span, the length skew and the
// term density within the span.  Widths are spans, never n: a row's columns
// all lie in its span, and for uniformly spread columns the span is n.
void product_row_features(const LView &L, const AView &A, float *F, hipStream_t s);

// One refill: Cx = values of kron(I, L) D A.  Lval / Ax / d are device arrays
// (d may be null = identity).  Launches on `s`, allocates nothing.
void product_numeric(const ProductPlan &P, const double *Lval, const double *d,
                     const double *Ax, double *Cx, hipStream_t s);

// K refills sharing operands: refill r reads d + r*ldd and writes Cx + r*ldc.
// opts: BATCH_VALUES_BOUND = the term operand pairs were gathered by
// product_bind_values for the current operand values (skip the gather);
// BATCH_D_KMAJOR = D is already k-major (D[k*ldd + r]).  ev_start / ev_stop
// (hipEvent_t, optional) are recorded around the main refill kernel.
enum : unsigned
{
    BATCH_VALUES_BOUND = 1u,
    BATCH_D_KMAJOR = 2u
};
void product_numeric_batched(const ProductPlan &P, const double *Lval, const double *D,
                             int64_t ldd, int K, const double *Ax, double *Cx, int64_t ldc,
                             hipStream_t s, const RefillCfg *cfg = nullptr, unsigned opts = 0,
                             void *ev_start = nullptr, void *ev_stop = nullptr,
                             const BatchScratch *ws = nullptr);
// Gather the term operand pairs for the interleaved refill (needs the reserved
// workspace; no-op otherwise).
void product_bind_values(const ProductPlan &P, const double *Lval, const double *Ax,
                         hipStream_t s);
RefillCfg default_refill_cfg(const ProductPlan &P, int K);
std::vector<RefillCfg> refill_candidates();

// Number of terms sum_r sum_{q in L_r} len(A_k); device pass, synchronises.
int64_t product_count_terms(const LView &L, const AView &A, int64_t *row_terms_out);

} // namespace sdl
