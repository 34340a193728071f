// Hessian assembly across expression nodes as one captured launch sequence
// (include/sdl_device.h, sdl_hessian_*): the caller-side batching of
// SparseDiffEngine's per-node fills.  Per Newton step a Lagrangian Hessian is
// refilled node by node -- BTDA_matrices_fill_values per chain-rule node
// (src/matmul_dispatchers.c:73-92), sum_n_matrices_fill_values over the nodes
// (src/sparse/matrix_sum.c:112-180), refresh_lower_triangular_coo for the
// solver (src/sparse/COO_matrix.c:92-98) -- and here the same steps run as
// device launches, eagerly or replayed from one HIP graph.
#include "../../include/sdl_device.h"
#include "btda_plan.h"
#include "sparse_ops.h"
#include "spgemm.h"

#include <algorithm>
#include <new>
#include <vector>

struct sdl_hessian
{
    int n = 0;
    std::vector<sdl_btda_plan *> nodes;
    std::vector<sdl::DBuf<double>> Cx;      // each node's output values
    std::vector<sdl::DBuf<int>> p32;        // each node's output row pointers, int32
    sdl_sum_plan *sum = nullptr;
    sdl_lower_coo *lower = nullptr;
    sdl_graph *graph = nullptr;
    // the node refills are independent: by default one launch covers every
    // node's tiles (streams == 0); else they fork onto these streams (default
    // priority) and join back before the sum, so a captured graph has one
    // branch per stream
    static constexpr int NAUX = 4;
    int streams = 0;  // 0: one launch; 1: in order on s; 2..4: forked
    hipStream_t aux[NAUX] = {nullptr};
    hipEvent_t fork = nullptr, join[NAUX] = {nullptr};
};

namespace {

hipStream_t as_stream(void *s) { return s ? (hipStream_t) s : sdl::stream(); }

void release(sdl_hessian *h)
{
    if (!h) return;
    for (int w = 0; w < sdl_hessian::NAUX; w++)
    {
        if (h->aux[w]) (void) hipStreamDestroy(h->aux[w]);
        if (h->join[w]) (void) hipEventDestroy(h->join[w]);
    }
    if (h->fork) (void) hipEventDestroy(h->fork);
    if (h->graph) sdl_graph_destroy(h->graph);
    if (h->lower) sdl_lower_coo_destroy(h->lower);
    if (h->sum) sdl_sum_plan_destroy(h->sum);
    for (sdl_btda_plan *p : h->nodes) sdl_btda_plan_destroy(p);
    delete h;
}

// The launches of one evaluation ordered after stream s's prior work: node
// refills forked over the auxiliary streams, joined back into s, then the sum
// and the lower-triangular export on s.  Allocates nothing, synchronises
// nothing; under capture of s the forked streams join the capture.
int eval_launches(sdl_hessian *h, const double *const *dAx, const double *const *dd,
                  double *dHx, double *dLx, hipStream_t s)
{
    const int k = (int) h->nodes.size();
    std::vector<const double *> xs(k);
    for (int j = 0; j < k; j++) xs[j] = h->Cx[j].p;
    if (h->streams == 0)
    {
        // every node's refill in one launch (A_j^T diag(d_j) A_j: B = A_j, so
        // both value arrays are dAx[j]), when every node has the tile layout
        std::vector<const sdl::ProductPlan *> plans(k);
        std::vector<double *> cx(k);
        for (int j = 0; j < k; j++)
        {
            plans[j] = &h->nodes[j]->P;
            cx[j] = h->Cx[j].p;
        }
        if (sdl::product_numeric_multi(plans.data(), k, dAx, dd, dAx, cx.data(), s))
        {
            if (sdl_sum_fill(h->sum, xs.data(), nullptr, dHx, s)) return 1;
            if (dLx && sdl_lower_coo_refresh(h->lower, dHx, dLx, s)) return 1;
            return 0;
        }
    }
    if (h->streams <= 1)
    {
        // every node refill in order on s
        for (int j = 0; j < k; j++)
        {
            if (sdl_btda_refill(h->nodes[j], dAx[j], dAx[j], dd ? dd[j] : nullptr, h->Cx[j].p, s))
                return 1;
            xs[j] = h->Cx[j].p;
        }
        if (sdl_sum_fill(h->sum, xs.data(), nullptr, dHx, s)) return 1;
        if (dLx && sdl_lower_coo_refresh(h->lower, dHx, dLx, s)) return 1;
        return 0;
    }
    const int naux = std::min(k, h->streams);
    SDL_HIP(hipEventRecord(h->fork, s));
    for (int w = 0; w < naux; w++) SDL_HIP(hipStreamWaitEvent(h->aux[w], h->fork, 0));
    for (int j = 0; j < k; j++)
    {
        if (sdl_btda_refill(h->nodes[j], dAx[j], dAx[j], dd ? dd[j] : nullptr, h->Cx[j].p,
                            h->aux[j % naux]))
            return 1;
        xs[j] = h->Cx[j].p;
    }
    for (int w = 0; w < naux; w++)
    {
        SDL_HIP(hipEventRecord(h->join[w], h->aux[w]));
        SDL_HIP(hipStreamWaitEvent(s, h->join[w], 0));
    }
    if (sdl_sum_fill(h->sum, xs.data(), nullptr, dHx, s)) return 1;
    if (dLx && sdl_lower_coo_refresh(h->lower, dHx, dLx, s)) return 1;
    return 0;
}

} // namespace

extern "C" {
#pragma GCC visibility push(default)

int sdl_hessian_create(int nnodes, int n, const int *m, const int *const *dAp,
                       const int *const *dAi, sdl_hessian **out)
{
    if (!out || nnodes < 1 || nnodes > 32 || n < 0 || !m || !dAp || !dAi)
    {
        sdl::set_error("sdl_hessian_create: invalid arguments (1 to 32 nodes)");
        return 1;
    }
    sdl_hessian *h = new (std::nothrow) sdl_hessian();
    if (!h) return 1;
    h->n = n;
    for (int w = 0; w < sdl_hessian::NAUX; w++)
    {
        SDL_HIP(hipStreamCreateWithFlags(&h->aux[w], hipStreamNonBlocking));
        SDL_HIP(hipEventCreateWithFlags(&h->join[w], hipEventDisableTiming));
    }
    SDL_HIP(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
    std::vector<const int *> P(nnodes), I(nnodes);
    std::vector<long long> nz(nnodes);
    for (int j = 0; j < nnodes; j++)
    {
        sdl_btda_plan *pl = nullptr;
        // A_j^T diag(d_j) A_j: B = A_j
        if (sdl_btda_plan_create(m[j], n, n, dAp[j], dAi[j], dAp[j], dAi[j], SDL_SYM_EXACT,
                                 &pl))
        {
            release(h);
            return 1;
        }
        h->nodes.push_back(pl);
        const long long nnz = sdl_btda_plan_nnz(pl);
        if (nnz > 0x7fffffffLL)
        {
            sdl::set_error("sdl_hessian_create: a node's output exceeds INT_MAX entries");
            release(h);
            return 1;
        }
        h->Cx.emplace_back(std::max<long long>(nnz, 1));
        h->p32.emplace_back((size_t) n + 1);
        sdl::narrow_i64((int64_t) n + 1, (const int64_t *) sdl_btda_plan_rowptr(pl),
                        h->p32.back().p, sdl::stream());
        P[j] = h->p32.back().p;
        I[j] = sdl_btda_plan_colind(pl);
        nz[j] = nnz;
    }
    if (sdl_sum_plan_create(nnodes, n, n, P.data(), I.data(), nz.data(), 0, &h->sum) ||
        sdl_lower_coo_create(n, sdl_sum_plan_rowptr(h->sum), sdl_sum_plan_colind(h->sum),
                             &h->lower))
    {
        release(h);
        return 1;
    }
    sdl::sync();
    *out = h;
    return 0;
}

void sdl_hessian_destroy(sdl_hessian *h) { release(h); }

int sdl_hessian_set_streams(sdl_hessian *h, int streams)
{
    if (!h || streams < 0 || streams > sdl_hessian::NAUX)
    {
        sdl::set_error("sdl_hessian_set_streams: 0 to 4 streams");
        return 1;
    }
    h->streams = streams;
    return 0;
}

long long sdl_hessian_nnz(const sdl_hessian *h) { return h ? sdl_sum_plan_nnz(h->sum) : -1; }
const long long *sdl_hessian_rowptr(const sdl_hessian *h)
{
    return h ? sdl_sum_plan_rowptr(h->sum) : nullptr;
}
const int *sdl_hessian_colind(const sdl_hessian *h)
{
    return h ? sdl_sum_plan_colind(h->sum) : nullptr;
}
long long sdl_hessian_lower_nnz(const sdl_hessian *h)
{
    return h ? sdl_lower_coo_nnz(h->lower) : -1;
}
const int *sdl_hessian_lower_rows(const sdl_hessian *h)
{
    return h ? sdl_lower_coo_rows(h->lower) : nullptr;
}
const int *sdl_hessian_lower_cols(const sdl_hessian *h)
{
    return h ? sdl_lower_coo_cols(h->lower) : nullptr;
}

int sdl_hessian_eval(sdl_hessian *h, const double *const *dAx, const double *const *dd,
                     double *dHx, double *dLx, void *stream)
{
    if (!h || !dAx || !dHx)
    {
        sdl::set_error("sdl_hessian_eval: invalid arguments");
        return 1;
    }
    return eval_launches(h, dAx, dd, dHx, dLx, as_stream(stream));
}

int sdl_hessian_capture(sdl_hessian *h, const double *const *dAx, const double *const *dd,
                        double *dHx, double *dLx, void *stream)
{
    if (!h || !dAx || !dHx)
    {
        sdl::set_error("sdl_hessian_capture: invalid arguments");
        return 1;
    }
    if (h->graph)
    {
        sdl_graph_destroy(h->graph);
        h->graph = nullptr;
    }
    if (sdl_graph_begin(stream)) return 1;
    const int rc = eval_launches(h, dAx, dd, dHx, dLx, as_stream(stream));
    sdl_graph *g = nullptr;
    const int rc2 = sdl_graph_end(stream, &g);
    if (rc || rc2)
    {
        if (g) sdl_graph_destroy(g);
        if (rc) sdl::set_error("sdl_hessian_capture: a node's launches failed under capture");
        return 1;
    }
    h->graph = g;
    return 0;
}

int sdl_hessian_replay(const sdl_hessian *h, void *stream)
{
    if (!h || !h->graph)
    {
        sdl::set_error("sdl_hessian_replay: nothing captured");
        return 1;
    }
    return sdl_graph_launch(h->graph, stream);
}

#pragma GCC visibility pop
} // extern "C"
