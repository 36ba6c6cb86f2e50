/*
 * layout.c -- LAPACK column-major <-> tile layout translation.
 *
 * plasma_omp_{d,s}{ge2desc,desc2ge,tr2desc,desc2tr} keep the reference's
 * signatures and validation (compute/zge2desc.c:27-68 and its siblings);
 * the per-tile lacpy tasks of pz{ge2desc,desc2ge,tr2desc,desc2tr}.c become
 * one coalesced kernel over all tiles.  The LAPACK array may be a device
 * pointer (translated in place on the GPU) or a host pointer (staged
 * through a device buffer; a copy back to the host happens only when the
 * sequence has not failed, as the reference's desc2* tasks are skipped on
 * a failed sequence).
 */
#include <string.h>

#include "internal.h"

/* Elements of the view covered by the region's tiles. */
static double region_elems(plasma_enum_t region, plasma_desc_t A)
{
    double e = 0;
    for (int n = 0; n < A.nt; n++)
        for (int m = 0; m < A.mt; m++) {
            if ((region == PlasmaLower && n > m) ||
                (region == PlasmaUpper && m > n))
                continue;
            e += (double)plasma_tile_mview(A, m) * plasma_tile_nview(A, n);
        }
    return e;
}

static int launch_l2t(plasma_enum_t region, const void *pA, int lda,
                      plasma_desc_t A, const int *st, hipStream_t s)
{
    int slot;
    plasma_context_t *ctx = plasma_context_self();
    plasma_prof_begin(s, PlasmaProfLayout, 0,
                      ctx->profile ? 2.0 * region_elems(region, A) *
                                     plasma_element_size(A.precision) : 0,
                      &slot);
    int err = plasma_kern_lapack2tile(region, pA, lda, A, st, s);
    plasma_prof_end(s, slot);
    return err;
}

static int launch_t2l(plasma_enum_t region, plasma_desc_t A, void *pA,
                      int lda, const int *st, hipStream_t s)
{
    int slot;
    plasma_context_t *ctx = plasma_context_self();
    plasma_prof_begin(s, PlasmaProfLayout, 0,
                      ctx->profile ? 2.0 * region_elems(region, A) *
                                     plasma_element_size(A.precision) : 0,
                      &slot);
    int err = plasma_kern_tile2lapack(region, A, pA, lda, st, s);
    plasma_prof_end(s, slot);
    return err;
}

void plasma_pxlapack2desc(plasma_enum_t region, const void *pA, int lda,
                          plasma_desc_t A, plasma_sequence_t *sequence,
                          plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    if (A.m == 0 || A.n == 0)
        return;
    if (plasma_is_device_ptr(pA)) {
        if (launch_l2t(region, pA, lda, A, st, ctx->stream))
            plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return;
    }
    size_t es = plasma_element_size(A.precision);
    void *buf = plasma_dev_alloc((size_t)A.m * A.n * es);
    if (buf == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    if (plasma_hip_check(hipMemcpy2DAsync(buf, (size_t)A.m * es, pA,
                                          (size_t)lda * es, (size_t)A.m * es,
                                          A.n, hipMemcpyHostToDevice,
                                          ctx->stream), "hipMemcpy2D") ||
        launch_l2t(region, buf, A.m, A, st, ctx->stream))
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
    plasma_dev_free(buf);
}

void plasma_pxdesc2lapack(plasma_enum_t region, plasma_desc_t A, void *pA,
                          int lda, plasma_sequence_t *sequence,
                          plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    if (A.m == 0 || A.n == 0)
        return;
    if (plasma_is_device_ptr(pA)) {
        if (launch_t2l(region, A, pA, lda, st, ctx->stream))
            plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return;
    }
    size_t es = plasma_element_size(A.precision);
    size_t pitch = (size_t)A.m * es;
    void *buf = plasma_dev_alloc(pitch * A.n);
    if (buf == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    /* entries outside the region keep the caller's values */
    if (region != PlasmaGeneral &&
        plasma_hip_check(hipMemcpy2DAsync(buf, pitch, pA, (size_t)lda * es,
                                          pitch, A.n, hipMemcpyHostToDevice,
                                          ctx->stream), "hipMemcpy2D")) {
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        plasma_dev_free(buf);
        return;
    }
    if (launch_t2l(region, A, buf, A.m, st, ctx->stream)) {
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        plasma_dev_free(buf);
        return;
    }
    if (plasma_sequence_sync(sequence) == PlasmaSuccess)
        plasma_hip_check(hipMemcpy2D(pA, (size_t)lda * es, buf, pitch, pitch,
                                     A.n, hipMemcpyDeviceToHost),
                         "hipMemcpy2D");
    plasma_dev_free(buf);
}

/* Argument checks shared by the four public translations
 * (reference compute/zge2desc.c:30-60). */
static int layout_args_ok(plasma_enum_t precision, const void *pA, int lda,
                          plasma_desc_t A, plasma_sequence_t *sequence,
                          plasma_request_t *request)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (plasma_desc_check(A) != PlasmaSuccess) {
        plasma_error("invalid A");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (A.precision != precision) {
        plasma_error("descriptor precision does not match the routine");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (sequence == NULL) {
        plasma_error("NULL sequence");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (request == NULL) {
        plasma_error("NULL request");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (pA == NULL || lda < imax(1, A.m)) {
        plasma_error("invalid LAPACK array");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    return 1;
}

#define DEFINE_LAYOUT(P, TYPE, PREC)                                          \
void plasma_omp_##P##ge2desc(TYPE *pA, int lda, plasma_desc_t A,              \
                             plasma_sequence_t *sequence,                     \
                             plasma_request_t *request)                       \
{                                                                             \
    if (!layout_args_ok(PREC, pA, lda, A, sequence, request)) return;        \
    plasma_pxlapack2desc(PlasmaGeneral, pA, lda, A, sequence, request);       \
}                                                                             \
void plasma_omp_##P##desc2ge(plasma_desc_t A, TYPE *pA, int lda,              \
                             plasma_sequence_t *sequence,                     \
                             plasma_request_t *request)                       \
{                                                                             \
    if (!layout_args_ok(PREC, pA, lda, A, sequence, request)) return;        \
    plasma_pxdesc2lapack(PlasmaGeneral, A, pA, lda, sequence, request);       \
}                                                                             \
void plasma_omp_##P##tr2desc(TYPE *pA, int lda, plasma_desc_t A,              \
                             plasma_sequence_t *sequence,                     \
                             plasma_request_t *request)                       \
{                                                                             \
    if (!layout_args_ok(PREC, pA, lda, A, sequence, request)) return;        \
    plasma_pxlapack2desc(A.type == PlasmaUpper ? PlasmaUpper : PlasmaLower,   \
                         pA, lda, A, sequence, request);                      \
}                                                                             \
void plasma_omp_##P##desc2tr(plasma_desc_t A, TYPE *pA, int lda,              \
                             plasma_sequence_t *sequence,                     \
                             plasma_request_t *request)                       \
{                                                                             \
    if (!layout_args_ok(PREC, pA, lda, A, sequence, request)) return;        \
    plasma_pxdesc2lapack(A.type == PlasmaUpper ? PlasmaUpper : PlasmaLower,   \
                         A, pA, lda, sequence, request);                      \
}

DEFINE_LAYOUT(d, double, PlasmaRealDouble)
DEFINE_LAYOUT(s, float, PlasmaRealFloat)
DEFINE_LAYOUT(z, plasma_complex64_t, PlasmaComplexDouble)
DEFINE_LAYOUT(c, plasma_complex32_t, PlasmaComplexFloat)
