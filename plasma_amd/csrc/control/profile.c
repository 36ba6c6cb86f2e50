/*
 * profile.c -- per-class kernel timing with HIP events (no reference
 * counterpart; the reference times whole calls with omp_get_wtime in its
 * tester, test/test_zposv.c:127-130).
 *
 * plasma_prof_begin / plasma_prof_end bracket one launch on the stream it
 * runs on.  Records are folded into per-class totals at query time, or when
 * the record table fills up.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static void prof_fold(plasma_context_t *ctx)
{
    for (int i = 0; i < ctx->nprof; i++) {
        plasma_prof_rec_t *r = &ctx->prof[i];
        float ms = 0.0f;
        if (hipEventSynchronize(r->end) == hipSuccess &&
            hipEventElapsedTime(&ms, r->beg, r->end) == hipSuccess) {
            ctx->prof_ms[r->cls] += ms;
            ctx->prof_flops[r->cls] += r->flops;
            ctx->prof_bytes[r->cls] += r->bytes;
            ctx->prof_count[r->cls] += 1;
        }
        else {
            (void)hipGetLastError();
        }
    }
    ctx->nprof = 0;
}

int plasma_profile_enable(int on)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL)
        return PlasmaErrorNotInitialized;
    if (on && ctx->prof == NULL) {
        ctx->prof = calloc(PLASMA_PROF_MAX, sizeof(plasma_prof_rec_t));
        if (ctx->prof == NULL)
            return PlasmaErrorOutOfMemory;
        for (int i = 0; i < PLASMA_PROF_MAX; i++) {
            if (hipEventCreate(&ctx->prof[i].beg) != hipSuccess ||
                hipEventCreate(&ctx->prof[i].end) != hipSuccess)
                return PlasmaErrorInternal;
        }
    }
    if (ctx->prof != NULL) {
        hipStreamSynchronize(ctx->panel_stream);
        hipStreamSynchronize(ctx->stream);
        ctx->nprof = 0;
    }
    memset(ctx->prof_ms, 0, sizeof(ctx->prof_ms));
    memset(ctx->prof_flops, 0, sizeof(ctx->prof_flops));
    memset(ctx->prof_bytes, 0, sizeof(ctx->prof_bytes));
    memset(ctx->prof_count, 0, sizeof(ctx->prof_count));
    ctx->profile = on ? 1 : 0;
    return PlasmaSuccess;
}

int plasma_profile_query(int cls, double *ms, double *flops, double *bytes,
                         long *launches)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL)
        return PlasmaErrorNotInitialized;
    if (cls < 0 || cls >= PlasmaProfClasses)
        return PlasmaErrorIllegalValue;
    if (ctx->prof != NULL)
        prof_fold(ctx);
    if (ms) *ms = ctx->prof_ms[cls];
    if (flops) *flops = ctx->prof_flops[cls];
    if (bytes) *bytes = ctx->prof_bytes[cls];
    if (launches) *launches = ctx->prof_count[cls];
    return PlasmaSuccess;
}

void plasma_prof_begin(hipStream_t s, int cls, double flops, double bytes,
                       int *slot)
{
    plasma_context_t *ctx = plasma_context_self();
    *slot = -1;
    if (ctx == NULL || !ctx->profile || ctx->prof == NULL)
        return;
    if (ctx->nprof == PLASMA_PROF_MAX)
        prof_fold(ctx);
    plasma_prof_rec_t *r = &ctx->prof[ctx->nprof];
    r->cls = cls;
    r->flops = flops;
    r->bytes = bytes;
    hipEventRecord(r->beg, s);
    *slot = ctx->nprof++;
}

void plasma_prof_end(hipStream_t s, int slot)
{
    plasma_context_t *ctx = plasma_context_self();
    if (slot < 0 || ctx == NULL || ctx->prof == NULL)
        return;
    hipEventRecord(ctx->prof[slot].end, s);
}
