// Internal declarations of the MI355X sparse-product library: error handling,
// the library stream, device buffers and the kernel launchers shared between
// the C-ABI front ends (drop-in structs, matrix handles, device-resident plans).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <utility>

namespace sdl {

// ---------------------------------------------------------------- errors ----
[[noreturn]] void fatal(const char *what, const char *file, int line);
void set_error(const char *fmt, ...);

#define SDL_HIP(expr)                                                              \
    do                                                                             \
    {                                                                              \
        hipError_t sdl_e_ = (expr);                                                \
        if (sdl_e_ != hipSuccess) ::sdl::fatal(hipGetErrorString(sdl_e_), __FILE__, __LINE__); \
    } while (0)

#define SDL_REQUIRE(cond, msg)                                                     \
    do                                                                             \
    {                                                                              \
        if (!(cond)) ::sdl::fatal(msg, __FILE__, __LINE__);                        \
    } while (0)

// FNV-1a, 64-bit (offset basis 0xcbf29ce484222325, prime 0x100000001b3): the
// checksum closing the library's saved files (predictor, tuning table).
inline uint64_t fnv1a64(const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *) p;
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; i++)
    {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

// The library's own stream (default priority, non-blocking w.r.t. stream 0).
hipStream_t stream();
void sync();

// ------------------------------------------------------------ allocation ----
void *dmalloc(size_t bytes);           // device memory; aborts on failure
void dfree(void *p);
void h2d(void *dst, const void *src, size_t bytes);   // ordered on stream()
void d2h(void *dst, const void *src, size_t bytes);   // ordered on stream(), synchronous
void d2d(void *dst, const void *src, size_t bytes);
void dzero(void *dst, size_t bytes);

// Owning device buffer.
template <typename T> struct DBuf
{
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o)
        {
            reset();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { reset(); }
    void alloc(size_t count)
    {
        reset();
        n = count;
        p = static_cast<T *>(dmalloc((count > 0 ? count : 1) * sizeof(T)));
    }
    void reset()
    {
        if (p) dfree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Device copy of host data (alloc phase).
template <typename T> DBuf<T> upload(const T *h, int64_t n)
{
    DBuf<T> d(n > 0 ? n : 0);
    if (n > 0) h2d(d.p, h, (size_t) n * sizeof(T));
    return d;
}

// ----------------------------------------------------------- primitives -----
// Exclusive scan of n int32 or int64 counts into int64 offsets; out has n+1
// entries and out[n] is the total.  Scratch is allocated internally (alloc
// phase only).  Returns the total (synchronises).
int64_t scan_i32_to_i64(const int *in, int64_t *out, int64_t n, bool want_total = true);
int64_t scan_i64(const int64_t *in, int64_t *out, int64_t n, bool want_total = true);

// Stable LSD radix sort of (key, value) pairs by the low `bits` bits of the key.
// keys/vals are sorted in place (ping-pong with internal scratch).
void radix_sort_u32(uint32_t *keys, uint32_t *vals, int64_t n, int bits);
void radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bits);

// iota: v[i] = i
void iota_u32(uint32_t *v, int64_t n);

} // namespace sdl
