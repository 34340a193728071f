// Library runtime: the library stream, device allocation with counters (the
// fill phase must not allocate; tests read the counters), error reporting.
#include "sdl_internal.h"
#include "../../include/sdl_device.h"

#include <atomic>
#include <cstdarg>
#include <cstring>
#include <mutex>

namespace sdl {

namespace {
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Plain streaming copy (16 bytes per lane), the calibration kernel for the
// HBM traffic counters: it moves exactly `bytes` in and `bytes` out.
__global__ void k_stream_copy(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, size_t n)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += st) dst[i] = src[i];
}

// Streaming writes (the write-ceiling calibration): constant values over n
// 16-byte (or 8-byte) words, grid-stride, non-temporal or ordinary stores.
template <bool NT>
__global__ void k_stream_write16(u32x4 *__restrict__ dst, size_t n)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t) gridDim.x * blockDim.x;
    u32x4 v = {0u, 0u, 0u, 0u};
    for (; i < n; i += st)
    {
        if (NT) __builtin_nontemporal_store(v, dst + i);
        else dst[i] = v;
    }
}

__global__ void k_stream_write8(double *__restrict__ dst, size_t n)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t) gridDim.x * blockDim.x;
    for (; i < n; i += st) __builtin_nontemporal_store(0.0, dst + i);
}

// The store pattern of the grouped batched refill without its reads: thread
// e of tile b / groups writes rows g*W .. min(K, g*W + W) - 1 of column e
// (g = b % groups), one non-temporal 8-byte store per row.
__global__ void __launch_bounds__(256)
    k_write_rows(double *__restrict__ C, int64_t ldc, int K, int64_t n, int groups, int W)
{
    const int g = (int) (blockIdx.x % (unsigned) groups);
    const int64_t e = (int64_t) (blockIdx.x / (unsigned) groups) * 256 + threadIdx.x;
    if (e >= n) return;
    const int r1 = min(K, g * W + W);
    for (int r = g * W; r < r1; r++) __builtin_nontemporal_store((double) r, C + (int64_t) r * ldc + e);
}

std::once_flag g_once;
hipStream_t g_stream = nullptr;
std::atomic<long long> g_dev_allocs{0};
std::atomic<long long> g_dev_frees{0};
thread_local char g_err[512] = {0};

void init_once()
{
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
    {
        std::fprintf(stderr,
                     "sparsediff_linalg_amd: no HIP device available (%s); this library "
                     "has no CPU path.\n",
                     e != hipSuccess ? hipGetErrorString(e) : "0 devices");
        std::abort();
    }
    // Default priority only: every stream of the library runs at priority 0.
    SDL_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
}
} // namespace

[[noreturn]] void fatal(const char *what, const char *file, int line)
{
    std::fprintf(stderr, "sparsediff_linalg_amd: fatal: %s (%s:%d)\n", what, file, line);
    std::fflush(stderr);
    std::abort();
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

hipStream_t stream()
{
    std::call_once(g_once, init_once);
    return g_stream;
}

void sync() { SDL_HIP(hipStreamSynchronize(stream())); }

void *dmalloc(size_t bytes)
{
    stream();
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
    {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "hipMalloc of %zu bytes failed: %s", bytes,
                      hipGetErrorString(e));
        fatal(msg, __FILE__, __LINE__);
    }
    g_dev_allocs++;
    return p;
}

void dfree(void *p)
{
    if (!p) return;
    // hipFree synchronises the device; fine: frees happen only in free paths.
    SDL_HIP(hipFree(p));
    g_dev_frees++;
}

void h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return;
    SDL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
}

void d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return;
    SDL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    SDL_HIP(hipStreamSynchronize(stream()));
}

void d2d(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return;
    SDL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
}

void dzero(void *dst, size_t bytes)
{
    if (bytes == 0) return;
    SDL_HIP(hipMemsetAsync(dst, 0, bytes, stream()));
}

// Sustained FP64 matrix rate: every wavefront runs NACC independent
// accumulators through iters rounds of v_mfma_f64_16x16x4_f64 on register
// operands (no memory traffic in the loop); one value per thread is stored.
typedef double dmf4 __attribute__((ext_vector_type(4)));
template <int NACC>
__global__ void __launch_bounds__(256) k_mfma_f64(int iters, double *__restrict__ out)
{
    dmf4 acc[NACC];
#pragma unroll
    for (int a = 0; a < NACC; a++) acc[a] = dmf4{0.0, 0.0, 0.0, 0.0};
    const double x = 1.0 + threadIdx.x * 1e-9, y = 1.0 - threadIdx.x * 1e-9;
    for (int it = 0; it < iters; it++)
    {
#pragma unroll
        for (int a = 0; a < NACC; a++)
            acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[a], 0, 0, 0);
    }
    double sum = 0.0;
#pragma unroll
    for (int a = 0; a < NACC; a++) sum = sum + acc[a][0] + acc[a][1] + acc[a][2] + acc[a][3];
    out[(size_t) blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

} // namespace sdl

extern "C" {
#pragma GCC visibility push(default)

int sdl_mfma_kernel(int nacc, int blocks, int iters, double *out, void *stream)
{
    if (!out || blocks < 1 || iters < 1 || (nacc != 1 && nacc != 2 && nacc != 4 && nacc != 8 &&
                                            nacc != 16))
    {
        sdl::set_error("sdl_mfma_kernel: nacc 1, 2, 4, 8 or 16, blocks and iters >= 1");
        return 1;
    }
    hipStream_t s = stream ? (hipStream_t) stream : sdl::stream();
    switch (nacc)
    {
    case 1: hipLaunchKernelGGL(sdl::k_mfma_f64<1>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 2: hipLaunchKernelGGL(sdl::k_mfma_f64<2>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 4: hipLaunchKernelGGL(sdl::k_mfma_f64<4>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 8: hipLaunchKernelGGL(sdl::k_mfma_f64<8>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    default: hipLaunchKernelGGL(sdl::k_mfma_f64<16>, dim3(blocks), dim3(256), 0, s, iters, out);
    }
    SDL_HIP(hipGetLastError());
    return 0;
}

const char *sdl_last_error(void) { return sdl::g_err; }

void sdl_init(void) { (void) sdl::stream(); }

void *sdl_library_stream(void) { return (void *) sdl::stream(); }

void sdl_synchronize(void) { sdl::sync(); }

long long sdl_device_alloc_count(void) { return sdl::g_dev_allocs.load(); }

long long sdl_device_free_count(void) { return sdl::g_dev_frees.load(); }

void *sdl_dev_alloc(size_t bytes) { return sdl::dmalloc(bytes); }

void sdl_dev_free(void *p) { sdl::dfree(p); }

int sdl_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    sdl::h2d(dst, src, bytes);
    sdl::sync();
    return 0;
}

int sdl_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    sdl::d2h(dst, src, bytes);
    return 0;
}

int sdl_memset_d(void *dst, int value, size_t bytes)
{
    SDL_HIP(hipMemsetAsync(dst, value, bytes, sdl::stream()));
    sdl::sync();
    return 0;
}

void *sdl_event_create(void)
{
    sdl::stream();
    hipEvent_t e;
    SDL_HIP(hipEventCreate(&e));
    return (void *) e;
}

void sdl_event_record(void *ev, void *stream)
{
    SDL_HIP(hipEventRecord((hipEvent_t) ev, stream ? (hipStream_t) stream : sdl::stream()));
}

float sdl_event_elapsed_ms(void *start, void *stop)
{
    SDL_HIP(hipEventSynchronize((hipEvent_t) stop));
    float ms = 0.f;
    SDL_HIP(hipEventElapsedTime(&ms, (hipEvent_t) start, (hipEvent_t) stop));
    return ms;
}

void sdl_event_destroy(void *ev)
{
    if (ev) SDL_HIP(hipEventDestroy((hipEvent_t) ev));
}

int sdl_copy_kernel(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes % 16 != 0 || ((uintptr_t) dst % 16) || ((uintptr_t) src % 16))
    {
        sdl::set_error("sdl_copy_kernel: 16-byte aligned sizes and pointers only");
        return 1;
    }
    size_t n = bytes / 16;
    hipLaunchKernelGGL(sdl::k_stream_copy, dim3(8192), dim3(256), 0,
                       stream ? (hipStream_t) stream : sdl::stream(), (const sdl::u32x4 *) src,
                       (sdl::u32x4 *) dst, n);
    SDL_HIP(hipGetLastError());
    return 0;
}

int sdl_write_kernel(void *dst, size_t bytes, int mode, void *stream)
{
    const size_t word = mode == 2 ? 8 : 16;
    if (mode < 0 || mode > 2 || bytes % word != 0 || ((uintptr_t) dst % word))
    {
        sdl::set_error("sdl_write_kernel: mode 0..2 with word-aligned size and pointer only");
        return 1;
    }
    hipStream_t s = stream ? (hipStream_t) stream : sdl::stream();
    if (mode == 2)
        hipLaunchKernelGGL(sdl::k_stream_write8, dim3(8192), dim3(256), 0, s, (double *) dst,
                           bytes / 8);
    else if (mode == 0)
        hipLaunchKernelGGL(sdl::k_stream_write16<true>, dim3(8192), dim3(256), 0, s,
                           (sdl::u32x4 *) dst, bytes / 16);
    else
        hipLaunchKernelGGL(sdl::k_stream_write16<false>, dim3(8192), dim3(256), 0, s,
                           (sdl::u32x4 *) dst, bytes / 16);
    SDL_HIP(hipGetLastError());
    return 0;
}

int sdl_write_rows_kernel(double *dC, long long ldc, int K, long long n, int groups, void *stream)
{
    if (!dC || K < 1 || n < 1 || ldc < n || groups < 1 || groups > 64)
    {
        sdl::set_error("sdl_write_rows_kernel: invalid arguments");
        return 1;
    }
    const int W = (K + groups - 1) / groups;
    const long long blocks = (n + 255) / 256 * groups;
    hipLaunchKernelGGL(sdl::k_write_rows, dim3((unsigned) blocks), dim3(256), 0,
                       stream ? (hipStream_t) stream : sdl::stream(), dC, (int64_t) ldc, K,
                       (int64_t) n, groups, W);
    SDL_HIP(hipGetLastError());
    return 0;
}

int sdl_device_name(char *buf, int len)
{
    sdl::stream();
    int dev = 0;
    SDL_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SDL_HIP(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, (size_t) len, "%s|%s|cu=%d|lds=%zu|l2=%d", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount, (size_t) prop.sharedMemPerBlock, prop.l2CacheSize);
    return 0;
}

#pragma GCC visibility pop
} // extern "C"
