// Batched B^T diag(d) A refill that allocates nothing: the workspace of the
// interleaved batched refill (operand pairs and the k-major copy of d) lives
// in a fixed device array of this library instead of a plan reservation
// (include/sdl_device.h, sdl_btda_refill_batched_static).
#include "../../include/sdl_device.h"
#include "btda_plan.h"
#include "spgemm.h"

#include <algorithm>
#include <mutex>

namespace {

// 32 MiB per device, held until the process ends.  Every call on one device
// shares it and leaves the last call's transposed d in it.
constexpr size_t SCRATCH_BYTES = size_t(32) << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {nullptr};

// Device address of `scratch` on the current device, looked up once per
// device.  In host code `scratch` names a host-side stand-in, so only this
// address may be handed to kernels.
double *scratch_ptr()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    {
        sdl::set_error("sdl_btda_refill_batched_static: no usable current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev])
    {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
        {
            sdl::set_error("sdl_btda_refill_batched_static: hipGetSymbolAddress failed: %s",
                           hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = static_cast<double *>(p);
    }
    return g_scratch[dev];
}

hipStream_t as_stream(void *s) { return s ? (hipStream_t) s : sdl::stream(); }

} // namespace

extern "C" {
#pragma GCC visibility push(default)

long long sdl_static_scratch_bytes(void) { return (long long) SCRATCH_BYTES; }

int sdl_btda_refill_batched_static(const sdl_btda_plan *plan, const double *dAx,
                                   const double *dBx, const double *dD, long long ldd, int K,
                                   double *dC, long long ldc, void *stream)
{
    if (!plan || K < 0 || ldd < plan->m || ldc < plan->P.nnz)
    {
        sdl::set_error("sdl_btda_refill_batched_static: invalid arguments");
        return 1;
    }
    if (K == 0 || plan->P.nnz == 0) return 0;
    const sdl::ProductPlan &P = plan->P;
    if (!P.pull || K == 1)
    {
        // no transposed d: the same launches as sdl_btda_refill_batched
        sdl::product_numeric_batched(P, dBx, dD, ldd, K, dAx, dC, ldc, as_stream(stream));
        return 0;
    }
    // the interleaved kernels (the plan's tuned one if it is interleaved)
    sdl::RefillCfg c(3, 8);
    if (P.tuned.variant >= 3) c = P.tuned;
    const int64_t need = sdl::batch_scratch_doubles(P, K) + 32;  // + alignment of the d part
    if (need * (int64_t) sizeof(double) > (int64_t) SCRATCH_BYTES)
    {
        sdl::set_error("sdl_btda_refill_batched_static: 2 * terms + m * K (K rounded up to 16) "
                       "doubles needed, %lld bytes, beyond the %zu-byte static scratch",
                       (long long) (need * (int64_t) sizeof(double)), SCRATCH_BYTES);
        return 1;
    }
    double *p = scratch_ptr();
    if (!p) return 1;
    sdl::BatchScratch ws;
    ws.tv = p;
    // d's k-major copy after the operand pairs, on a 256-byte boundary
    ws.dT = p + (2 * std::max<int64_t>(P.terms, 1) + 31) / 32 * 32;
    ws.ld = ((int64_t) K + 15) / 16 * 16;
    sdl::product_numeric_batched(P, dBx, dD, ldd, K, dAx, dC, ldc, as_stream(stream), &c, 0,
                                 nullptr, nullptr, &ws);
    return 0;
}

#pragma GCC visibility pop
} // extern "C"
