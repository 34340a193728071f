// Device-resident C ABI (include/sdl_device.h): B^T diag(d) A plans on HBM
// operands, single and batched refills.
#include "../../include/sdl_device.h"
#include "btda_plan.h"
#include "mlp.h"
#include "sparse_ops.h"
#include "spgemm.h"

#include <climits>
#include <new>
#include <vector>


namespace {
hipStream_t as_stream(void *s) { return s ? (hipStream_t) s : sdl::stream(); }

int nnz_of(const int *dp, int m)
{
    int v = 0;
    sdl::d2h(&v, dp + m, sizeof(int));
    return v;
}
} // namespace

extern "C" {
#pragma GCC visibility push(default)

int sdl_btda_plan_create(int m, int nA, int nB, const int *dAp, const int *dAi,
                         const int *dBp, const int *dBi, int mode, sdl_btda_plan **out)
{
    const int sym = mode & 0xff;
    return sdl_btda_plan_create_mode(m, nA, nB, dAp, dAi, dBp, dBi,
                                     sym == SDL_SYM_PREDICTED ? sdl::SymMode::Predicted
                                                              : sdl::SymMode::Exact,
                                     sdl::g_row_predictor, out, (mode & SDL_PLAN_PULL) != 0,
                                     (mode & SDL_PLAN_RANK_REFILL) != 0);
}
#pragma GCC visibility pop
} // extern "C"

int sdl_btda_plan_create_mode(int m, int nA, int nB, const int *dAp, const int *dAi,
                              const int *dBp, const int *dBi, sdl::SymMode mode,
                              const sdl::RowPredictor *pred, sdl_btda_plan **out,
                              bool force_pull, bool rank_refill)
{
    if (!out || m < 0 || nA < 0 || nB < 0 || !dAp || !dBp)
    {
        sdl::set_error("sdl_btda_plan_create: invalid arguments");
        return 1;
    }
    sdl_btda_plan *pl = new (std::nothrow) sdl_btda_plan();
    if (!pl)
    {
        sdl::set_error("sdl_btda_plan_create: out of host memory");
        return 1;
    }
    pl->m = m;
    pl->nA = nA;
    pl->nB = nB;
    int nnzA = nnz_of(dAp, m), nnzB = nnz_of(dBp, m);
    pl->Ap.alloc((size_t) m + 1);
    pl->Ai.alloc(nnzA);
    sdl::d2d(pl->Ap.p, dAp, ((size_t) m + 1) * sizeof(int));
    sdl::d2d(pl->Ai.p, dAi, (size_t) nnzA * sizeof(int));
    pl->Lp.alloc((size_t) nB + 1);
    pl->Li.alloc(nnzB);
    pl->Lperm.alloc(nnzB);
    sdl::transpose_structure(m, nB, nnzB, dBp, dBi, pl->Lp.p, pl->Li.p, pl->Lperm.p);
    sdl::LView L;
    L.Lm = nB;
    L.Lp = pl->Lp.p;
    L.Li = pl->Li.p;
    L.Lperm = pl->Lperm.p;
    L.blocks = 1;
    sdl::AView A;
    A.Am = m;
    A.An = nA;
    A.Ap = pl->Ap.p;
    A.Ai = pl->Ai.p;
    pl->P.rank_refill = rank_refill;
    sdl::product_plan(pl->P, L, A, mode, pred, true, force_pull);
    *out = pl;
    return 0;
}

extern "C" {
#pragma GCC visibility push(default)

void sdl_btda_plan_destroy(sdl_btda_plan *plan)
{
    if (!plan) return;
    sdl::sync();
    delete plan;
}

long long sdl_btda_plan_nnz(const sdl_btda_plan *plan) { return plan ? plan->P.nnz : -1; }
long long sdl_btda_plan_terms(const sdl_btda_plan *plan) { return plan ? plan->P.terms : -1; }
int sdl_btda_plan_is_pull(const sdl_btda_plan *plan) { return plan ? (int) plan->P.pull : 0; }

const long long *sdl_btda_plan_rowptr(const sdl_btda_plan *plan)
{
    return plan ? (const long long *) plan->P.Cp.p : nullptr;
}

const int *sdl_btda_plan_colind(const sdl_btda_plan *plan) { return plan ? plan->P.Ci.p : nullptr; }

int sdl_btda_plan_structure(const sdl_btda_plan *plan, int *hp, int *hi)
{
    if (!plan) return 1;
    if (plan->P.nnz > INT_MAX)
    {
        sdl::set_error("sdl_btda_plan_structure: nnz %lld exceeds INT_MAX", (long long) plan->P.nnz);
        return 1;
    }
    sdl::DBuf<int> p32(plan->P.m_out + 1);
    sdl::narrow_i64(plan->P.m_out + 1, plan->P.Cp.p, p32.p, sdl::stream());
    sdl::d2h(hp, p32.p, (size_t) (plan->P.m_out + 1) * sizeof(int));
    sdl::d2h(hi, plan->P.Ci.p, (size_t) plan->P.nnz * sizeof(int));
    return 0;
}

int sdl_btda_plan_stats(const sdl_btda_plan *plan, long long *o)
{
    if (!plan || !o) return 1;
    const sdl::SymStats &s = plan->P.stats;
    o[0] = s.rows;
    o[1] = s.terms;
    o[2] = s.nnz;
    o[3] = s.hash_rows;
    o[4] = s.bitmap_rows;
    o[5] = s.global_rows;
    o[6] = s.overflow_rows;
    o[7] = s.predicted_rows;
    o[8] = s.count_hash_rows;
    o[9] = s.count_bitmap_rows;
    o[10] = s.count_global_rows;
    return 0;
}

int sdl_btda_plan_window_rows(const sdl_btda_plan *plan, long long *o)
{
    if (!plan || !o) return 1;
    const sdl::ProductPlan &P = plan->P;
    o[0] = P.stats.count_window_rows;
    o[1] = P.stats.window_rows;
    o[2] = P.pull ? 0 : P.rwin_off[7] - P.rwin_off[1];
    o[3] = P.pull ? 0 : P.rwide_off[3];
    return 0;
}

int sdl_symbolic_window(int mode)
{
    if (mode < -1 || mode > 1) return -1;
    const int prev = sdl::g_sym_window;
    if (mode >= 0) sdl::g_sym_window = mode;
    return prev;
}

int sdl_symbolic_lane_groups(int mode)
{
    if (mode < -1 || mode > 1) return -1;
    const int prev = sdl::g_sym_groups;
    if (mode >= 0) sdl::g_sym_groups = mode;
    return prev;
}

int sdl_rowacc_window(int mode)
{
    if (mode < -1 || mode > 1) return -1;
    const int prev = sdl::g_rowacc_window;
    if (mode >= 0) sdl::g_rowacc_window = mode;
    return prev;
}

int sdl_symbolic_window_lanes(int lanes)
{
    if (lanes < -1 || lanes > 64 || (lanes > 0 && (lanes < 4 || (lanes & (lanes - 1)))))
        return -1;
    const int prev = sdl::g_sym_window_lanes;
    if (lanes >= 0) sdl::g_sym_window_lanes = lanes;
    return prev;
}

int sdl_btda_refill(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                    const double *dd, double *dCx, void *stream)
{
    if (!plan) return 1;
    sdl::product_numeric(plan->P, dBx, dd, dAx, dCx, as_stream(stream));
    return 0;
}

int sdl_btda_refill_range(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                          const double *dd, double *dCx, long long e0, long long e1, void *stream)
{
    if (!plan || !plan->P.pull || e0 < 0 || e1 < e0 || e1 > plan->P.nnz)
    {
        sdl::set_error("sdl_btda_refill_range: invalid range or the plan has no pull schedule");
        return 1;
    }
    sdl::product_numeric_range(plan->P, dBx, dd, dAx, dCx, e0, e1, as_stream(stream));
    return 0;
}

int sdl_btda_refill_batched(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                            const double *dD, long long ldd, int K, double *dC, long long ldc,
                            void *stream)
{
    if (!plan || K < 0 || ldd < plan->m || ldc < plan->P.nnz)
    {
        sdl::set_error("sdl_btda_refill_batched: invalid arguments");
        return 1;
    }
    sdl::product_numeric_batched(plan->P, dBx, dD, ldd, K, dAx, dC, ldc, as_stream(stream));
    return 0;
}

int sdl_btda_refill_batched_cfg(const sdl_btda_plan *plan, const double *dAx,
                                const double *dBx, const double *dD, long long ldd, int K,
                                double *dC, long long ldc, const sdl_refill_cfg *cfg,
                                void *stream)
{
    if (!plan || !cfg || K < 0 || ldd < plan->m || ldc < plan->P.nnz)
    {
        sdl::set_error("sdl_btda_refill_batched_cfg: invalid arguments");
        return 1;
    }
    sdl::RefillCfg c;
    c.variant = cfg->variant;
    c.rchunk = cfg->rchunk;
    sdl::product_numeric_batched(plan->P, dBx, dD, ldd, K, dAx, dC, ldc, as_stream(stream), &c);
    return 0;
}

int sdl_refill_candidates(sdl_refill_cfg *out, int max)
{
    std::vector<sdl::RefillCfg> v = sdl::refill_candidates();
    int n = 0;
    for (auto &c : v)
    {
        if (n >= max) break;
        out[n].variant = c.variant;
        out[n].rchunk = c.rchunk;
        n++;
    }
    return (int) v.size();
}

int sdl_btda_autotune(sdl_btda_plan *plan, const double *dAx, const double *dBx,
                      const double *dD, long long ldd, int K, double *dC, long long ldc,
                      int reps, sdl_refill_cfg *best, float *best_us)
{
    if (!plan || K <= 0 || ldd < plan->m || ldc < plan->P.nnz)
    {
        sdl::set_error("sdl_btda_autotune: invalid arguments");
        return 1;
    }
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    SDL_HIP(hipEventCreate(&e0));
    SDL_HIP(hipEventCreate(&e1));
    hipStream_t s = sdl::stream();
    float best_t = 1e30f;
    sdl::RefillCfg best_c = sdl::default_refill_cfg(plan->P, K);
    for (const sdl::RefillCfg &c : sdl::refill_candidates())
    {
        // without an operand-pair array (tile reservation) the pair-reading
        // variants run as the tile form: time that once, as itself
        if (!plan->P.tv.p && c.variant >= 3 && c.variant != 7) continue;
        for (int w = 0; w < 2; w++)
            sdl::product_numeric_batched(plan->P, dBx, dD, ldd, K, dAx, dC, ldc, s, &c);
        SDL_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++)
            sdl::product_numeric_batched(plan->P, dBx, dD, ldd, K, dAx, dC, ldc, s, &c);
        SDL_HIP(hipEventRecord(e1, s));
        SDL_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SDL_HIP(hipEventElapsedTime(&ms, e0, e1));
        float us = ms * 1000.f / reps;
        if (us < best_t)
        {
            best_t = us;
            best_c = c;
        }
    }
    SDL_HIP(hipEventDestroy(e0));
    SDL_HIP(hipEventDestroy(e1));
    plan->P.tuned = best_c;
    {
        sdl::SymTune keep = plan->P.tune;
        sdl::tune_lookup(plan->P.sig, &keep, nullptr);
        sdl::tune_record(plan->P.sig, &keep, &best_c);
    }
    if (best)
    {
        best->variant = best_c.variant;
        best->rchunk = best_c.rchunk;
    }
    if (best_us) *best_us = best_t;
    return 0;
}

int sdl_btda_bind_values(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                         void *stream)
{
    if (!plan) return 1;
    if (plan->P.batch_cap == 0)
    {
        sdl::set_error("sdl_btda_bind_values: no batch workspace reserved");
        return 1;
    }
    sdl::product_bind_values(plan->P, dBx, dAx, as_stream(stream));
    return 0;
}

int sdl_btda_refill_batched_ex(const sdl_btda_plan *plan, const double *dAx, const double *dBx,
                               const double *dD, long long ldd, int K, double *dC,
                               long long ldc, unsigned flags, void *ev_start, void *ev_stop,
                               void *stream)
{
    if (!plan || K < 0 || ldc < plan->P.nnz)
    {
        sdl::set_error("sdl_btda_refill_batched_ex: invalid arguments");
        return 1;
    }
    // a k-major D is read in whole runs of 16 refills per row, two values
    // (16 bytes) per load
    if ((flags & SDL_BATCH_D_KMAJOR)
            ? (ldd < ((long long) K + 15) / 16 * 16 || (ldd & 1) || ((uintptr_t) dD & 15))
            : ldd < plan->m)
    {
        sdl::set_error("sdl_btda_refill_batched_ex: ldd too small for the layout");
        return 1;
    }
    if ((flags & (SDL_BATCH_D_KMAJOR | SDL_BATCH_VALUES_BOUND)) &&
        (K > plan->P.batch_cap || !plan->P.pull))
    {
        sdl::set_error("sdl_btda_refill_batched_ex: K-major / bound refills need a pull plan "
                       "with a reservation >= K");
        return 1;
    }
    unsigned o = 0;
    if (flags & SDL_BATCH_VALUES_BOUND) o |= sdl::BATCH_VALUES_BOUND;
    if (flags & SDL_BATCH_D_KMAJOR) o |= sdl::BATCH_D_KMAJOR;
    sdl::product_numeric_batched(plan->P, dBx, dD, ldd, K, dAx, dC, ldc, as_stream(stream),
                                 nullptr, o, ev_start, ev_stop);
    return 0;
}

int sdl_btda_plan_reserve(sdl_btda_plan *plan, int Kmax)
{
    if (!plan || Kmax < 0) return 1;
    sdl::product_reserve_batch(plan->P, Kmax);
    sdl::sync();
    return 0;
}

int sdl_btda_plan_reserve_tile(sdl_btda_plan *plan, int Kmax)
{
    if (!plan || Kmax < 0) return 1;
    if (!plan->P.pull || !plan->P.tq.p)
    {
        sdl::set_error("sdl_btda_plan_reserve_tile: needs a pull schedule over a single-block L");
        return 1;
    }
    sdl::product_reserve_batch(plan->P, Kmax, true);
    sdl::sync();
    return 0;
}

int sdl_btda_plan_schedule(const sdl_btda_plan *plan, long long *out5)
{
    if (!plan || !out5) return 1;
    const sdl::ProductPlan &P = plan->P;
    out5[0] = P.pull ? 1 : 0;
    out5[1] = P.pull ? 0 : P.rbin_off[2] - P.rbin_off[1] - P.rwide_off[3];
    out5[2] = P.rankpos ? 1 : 0;
    out5[3] = P.rankpos ? P.nmulti : 0;
    int ns = 0;
    if (P.rankpos) sdl::d2h(&ns, P.nspill.p, sizeof ns);
    out5[4] = ns;
    return 0;
}

int sdl_btda_set_refill_cfg(sdl_btda_plan *plan, const sdl_refill_cfg *cfg)
{
    if (!plan) return 1;
    if (!cfg)
    {
        plan->P.tuned = sdl::RefillCfg();
        return 0;
    }
    plan->P.tuned.variant = cfg->variant;
    plan->P.tuned.rchunk = cfg->rchunk;
    return 0;
}

int sdl_btda_plan_single_refill(sdl_btda_plan *plan, int kind)
{
    if (!plan) return -1;
    sdl::ProductPlan &P = plan->P;
    const int prev = P.tile_refill ? 1 : 0;
    if (kind == 0) P.tile_refill = false;
    else if (kind == 1)
    {
        if (!P.pull || !P.tq.p)
        {
            sdl::set_error("sdl_btda_plan_single_refill: no tile refill for this plan "
                           "(row-accumulator schedule or block-diagonal L)");
            return -1;
        }
        P.tile_refill = true;
    }
    else if (kind != -1)
    {
        sdl::set_error("sdl_btda_plan_single_refill: kind must be -1, 0 or 1");
        return -1;
    }
    return prev;
}

// ---- COO coalescing on device-resident triplets ----
struct sdl_coo_plan
{
    sdl::CoalescePlan P;
};

int sdl_coo_plan_create(long long n_items, const int *drows, const int *dcols, int m, int n,
                        sdl_coo_plan **out)
{
    if (!out || n_items < 0 || m < 0 || n < 0 || (n_items > 0 && (!drows || !dcols)) ||
        n_items >= (1ll << 32))
    {
        sdl::set_error("sdl_coo_plan_create: invalid arguments");
        return 1;
    }
    sdl_coo_plan *p = new (std::nothrow) sdl_coo_plan();
    if (!p) return 1;
    sdl::coalesce_plan(p->P, n_items, drows, dcols, m, n);
    *out = p;
    return 0;
}

long long sdl_coo_plan_nnz(const sdl_coo_plan *p) { return p ? p->P.nnz : -1; }
const long long *sdl_coo_plan_rowptr(const sdl_coo_plan *p)
{
    return p ? (const long long *) p->P.Cp.p : nullptr;
}
const int *sdl_coo_plan_colind(const sdl_coo_plan *p) { return p ? p->P.Ci.p : nullptr; }

int sdl_coo_fill(const sdl_coo_plan *p, const double *dx, double *dCx, void *stream)
{
    if (!p) return 1;
    sdl::coalesce_fill(p->P, dx, dCx, as_stream(stream));
    return 0;
}

void sdl_coo_plan_destroy(sdl_coo_plan *p)
{
    if (!p) return;
    sdl::sync();
    delete p;
}

// ---- stream capture into HIP graphs ----
struct sdl_graph
{
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int sdl_graph_begin(void *stream)
{
    hipError_t e = hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeRelaxed);
    if (e != hipSuccess)
    {
        sdl::set_error("sdl_graph_begin: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}

int sdl_graph_end(void *stream, sdl_graph **out)
{
    if (!out) return 1;
    sdl_graph *g = new (std::nothrow) sdl_graph();
    if (!g) return 1;
    hipError_t e = hipStreamEndCapture(as_stream(stream), &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess)
    {
        sdl::set_error("sdl_graph_end: %s", hipGetErrorString(e));
        if (g->graph) (void) hipGraphDestroy(g->graph);
        delete g;
        return 1;
    }
    *out = g;
    return 0;
}

int sdl_graph_launch(const sdl_graph *g, void *stream)
{
    if (!g) return 1;
    hipError_t e = hipGraphLaunch(g->exec, as_stream(stream));
    if (e != hipSuccess)
    {
        sdl::set_error("sdl_graph_launch: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}

void sdl_graph_destroy(sdl_graph *g)
{
    if (!g) return;
    sdl::sync();
    if (g->exec) (void) hipGraphExecDestroy(g->exec);
    if (g->graph) (void) hipGraphDestroy(g->graph);
    delete g;
}

#pragma GCC visibility pop
} // extern "C"
