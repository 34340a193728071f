/* Host allocation wrappers of the sparsediff_linalg API, as the reference's
 * include/sparsediff_linalg/tracked_alloc.h:18-136 declares them.
 *
 * Compiled with -DSP_TRACK_MEMORY, sp_malloc / sp_calloc / sp_realloc /
 * sp_free keep two counters of host bytes (the live total and its high-water
 * mark since the last reset, block sizes as malloc_usable_size reports them)
 * that sp_memory_stats reports and sp_memory_reset_peak resets.  The library
 * (libsdl_amd.so) defines the counters and counts its own host allocations in
 * them: the arrays of every CSR / CSC / COO struct and block buffer it creates
 * and frees.  Device memory is not host memory and is not counted here
 * (sdl_device_alloc_count in sdl_device.h counts device allocations).
 * Without SP_TRACK_MEMORY the wrappers are libc's and nothing is counted by
 * the caller's code. */
#ifndef TRACKED_ALLOC_H
#define TRACKED_ALLOC_H

#include <stddef.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef SP_TRACK_MEMORY

#include <malloc.h>

/* The counters behind the inline wrappers (defined by the library; visible
   here so the wrappers can inline).  Read them through sp_memory_stats. */
extern size_t g_allocated_bytes; /* live host bytes */
extern size_t g_peak_bytes;      /* largest live total since the last reset */

/* Live bytes and high-water mark; either pointer may be NULL. */
void sp_memory_stats(size_t *live_bytes, size_t *peak_bytes);
/* High-water mark := current live bytes. */
void sp_memory_reset_peak(void);

static inline void sp_count_add_(void *p)
{
    if (!p) return;
    g_allocated_bytes += malloc_usable_size(p);
    if (g_allocated_bytes > g_peak_bytes) g_peak_bytes = g_allocated_bytes;
}

static inline void *sp_malloc(size_t size)
{
    void *p = malloc(size);
    sp_count_add_(p);
    return p;
}

static inline void *sp_calloc(size_t count, size_t size)
{
    void *p = calloc(count, size);
    sp_count_add_(p);
    return p;
}

static inline void sp_free(void *ptr)
{
    if (!ptr) return;
    g_allocated_bytes -= malloc_usable_size(ptr);
    free(ptr);
}

static inline void *sp_realloc(void *ptr, size_t size)
{
    size_t before = ptr ? malloc_usable_size(ptr) : 0;
    void *p = realloc(ptr, size);
    if (p) /* on failure the old block stays live and counted */
    {
        g_allocated_bytes -= before;
        sp_count_add_(p);
    }
    return p;
}

#else /* !SP_TRACK_MEMORY */

static inline void *sp_malloc(size_t size) { return malloc(size); }
static inline void *sp_calloc(size_t count, size_t size) { return calloc(count, size); }
static inline void sp_free(void *ptr) { free(ptr); }
static inline void *sp_realloc(void *ptr, size_t size) { return realloc(ptr, size); }

#endif /* SP_TRACK_MEMORY */

#ifdef __cplusplus
}
#endif

#endif /* TRACKED_ALLOC_H */
