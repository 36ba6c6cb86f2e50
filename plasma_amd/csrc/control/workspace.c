/*
 * workspace.c -- device memory for tile storage and staging buffers.
 *
 * The reference mallocs tile storage on every driver call (control/
 * descriptor.c:39-45, :107-113) and keeps per-thread host scratch
 * (control/workspace.c).  On the GPU a hipMalloc of gigabytes per call would
 * dominate short solves, so freed blocks stay cached and are handed out
 * again best-fit.  All users of pool memory enqueue their work on the
 * library's streams, which are joined before any driver returns, so a
 * block returned here can be reused by the next allocation in stream order.
 * PLASMA_POOL=0 in the environment disables caching.
 */
#include <stdlib.h>

#include "internal.h"

static int pool_enabled(void)
{
    static int flag = -1;
    if (flag < 0) {
        const char *e = getenv("PLASMA_POOL");
        flag = (e == NULL || atoi(e) != 0);
    }
    return flag;
}

static void pool_trim(plasma_context_t *ctx)
{
    hipStreamSynchronize(ctx->panel_stream);
    hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < PLASMA_POOL_ENTRIES; i++) {
        plasma_pool_entry_t *e = &ctx->pool[i];
        if (e->ptr != NULL && !e->in_use) {
            hipFree(e->ptr);
            e->ptr = NULL;
            e->bytes = 0;
        }
    }
}

void *plasma_dev_alloc(size_t bytes)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL)
        return NULL;
    if (bytes == 0)
        bytes = 256;
    /* round to 2 MiB so nearby sizes share blocks */
    bytes = (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);

    if (pool_enabled()) {
        int best = -1;
        for (int i = 0; i < PLASMA_POOL_ENTRIES; i++) {
            plasma_pool_entry_t *e = &ctx->pool[i];
            if (e->ptr != NULL && !e->in_use && e->bytes >= bytes &&
                e->bytes <= 2 * bytes &&
                (best < 0 || e->bytes < ctx->pool[best].bytes))
                best = i;
        }
        if (best >= 0) {
            ctx->pool[best].in_use = 1;
            return ctx->pool[best].ptr;
        }
    }

    void *p = NULL;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        pool_trim(ctx);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return NULL;
        }
    }
    if (pool_enabled()) {
        for (int i = 0; i < PLASMA_POOL_ENTRIES; i++) {
            plasma_pool_entry_t *e = &ctx->pool[i];
            if (e->ptr == NULL) {
                e->ptr = p;
                e->bytes = bytes;
                e->in_use = 1;
                return p;
            }
        }
    }
    return p;  /* untracked: freed directly */
}

void plasma_dev_free(void *ptr)
{
    if (ptr == NULL)
        return;
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL)
        return;
    for (int i = 0; i < PLASMA_POOL_ENTRIES; i++) {
        if (ctx->pool[i].ptr == ptr) {
            ctx->pool[i].in_use = 0;
            return;
        }
    }
    /* not cached: release after the work queued so far */
    hipStreamSynchronize(ctx->panel_stream);
    hipStreamSynchronize(ctx->stream);
    hipFree(ptr);
}

void plasma_pool_release_all(plasma_context_t *ctx)
{
    for (int i = 0; i < PLASMA_POOL_ENTRIES; i++) {
        if (ctx->pool[i].ptr != NULL)
            hipFree(ctx->pool[i].ptr);
        ctx->pool[i].ptr = NULL;
        ctx->pool[i].bytes = 0;
        ctx->pool[i].in_use = 0;
    }
}

int plasma_is_device_ptr(const void *p)
{
    hipPointerAttribute_t attr;
    if (p == NULL)
        return 0;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return attr.type == hipMemoryTypeDevice;
}
