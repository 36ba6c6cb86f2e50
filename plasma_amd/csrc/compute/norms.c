/*
 * norms.c -- plasma_{d,s}lansy, plasma_{d,s}lange, plasma_{d,s}amax and
 * their asynchronous tile forms plasma_omp_*.
 *
 * Signatures, argument checks and return conventions follow the reference
 * (compute/zlansy.c:81-300, compute/zlange.c:76-290, compute/dzamax.c:21-
 * 170): the synchronous drivers return the norm (a negative value is an
 * argument code), translate the LAPACK array into tiles and call the tile
 * form; the tile forms fail the sequence with PlasmaErrorIllegalValue on a
 * bad argument.  The kernels are in kernels/norms.hip.
 *
 * Extensions of the tile forms (device execution): `work` is device memory
 * of at least plasma_norm_work_doubles(A) doubles, or NULL for library
 * workspace; `value` / `values` may be host or device memory and hold the
 * result once the sequence is synchronized (plasma_sequence_sync).  Views
 * must start on a tile boundary.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int valid_norm(plasma_enum_t norm)
{
    return norm == PlasmaMaxNorm || norm == PlasmaOneNorm ||
           norm == PlasmaInfNorm || norm == PlasmaFrobeniusNorm;
}

/* Result of `count` doubles from device scratch into the caller's array
 * (device or host), in stream order; then the scratch goes back to the
 * pool (later users of the block come after the copy on the same stream). */
static void deliver(plasma_context_t *ctx, void *dst, const double *dsrc,
                    size_t count, plasma_enum_t precision,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (precision == PlasmaRealDouble) {
        if (plasma_hip_check(hipMemcpyAsync(dst, dsrc, count * sizeof(double),
                                            hipMemcpyDefault, ctx->stream),
                             "hipMemcpyAsync(norm)"))
            plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return;
    }
    /* single precision results: converted on the host after a sync */
    double *h = malloc(count * sizeof(double));
    float *f = malloc(count * sizeof(float));
    if (h == NULL || f == NULL ||
        plasma_hip_check(hipMemcpyAsync(h, dsrc, count * sizeof(double),
                                        hipMemcpyDeviceToHost, ctx->stream),
                         "hipMemcpyAsync(norm)") ||
        plasma_hip_check(hipStreamSynchronize(ctx->stream), "sync")) {
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
    }
    else {
        for (size_t i = 0; i < count; i++)
            f[i] = (float)h[i];
        if (plasma_hip_check(hipMemcpy(dst, f, count * sizeof(float),
                                       hipMemcpyDefault),
                             "hipMemcpy(norm)"))
            plasma_request_fail(sequence, request, PlasmaErrorInternal);
    }
    free(h);
    free(f);
}

static int tile_aligned(plasma_desc_t A)
{
    return A.i % A.mb == 0 && A.j % A.nb == 0;
}

/* norm of A (uplo PlasmaGeneral: lange) into the caller's value, a double
 * (vprec PlasmaRealDouble) or a float (PlasmaRealFloat) */
static void omp_norm(plasma_enum_t norm, plasma_enum_t uplo, plasma_desc_t A,
                     double *work, void *value, plasma_enum_t vprec,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!valid_norm(norm)) {
        plasma_error("illegal value of norm");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (uplo != PlasmaGeneral && uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (plasma_desc_check(A) != PlasmaSuccess) {
        plasma_error("invalid descriptor A");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (sequence == NULL) {
        plasma_error("NULL sequence");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (request == NULL) {
        plasma_error("NULL request");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (value == NULL) {
        plasma_error("NULL value");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!tile_aligned(A)) {
        plasma_error("views must start on a tile boundary");
        plasma_request_fail(sequence, request, PlasmaErrorNotSupported);
        return;
    }
    if (sequence->status != PlasmaSuccess)
        return;
    int *st = plasma_sequence_dstatus(sequence, request);
    /* quick return (reference zlansy.c: A.m == 0, zlange.c: min(m, n) == 0):
     * the result is 0, written in stream order like any other result */
    const int empty = (uplo == PlasmaGeneral) ? imin(A.m, A.n) == 0 : A.m == 0;
    const size_t wd = empty ? 1 : plasma_norm_work_doubles(A);
    double *own = NULL;
    if (empty || work == NULL) {
        own = plasma_dev_alloc((wd + 1) * sizeof(double));
        if (own == NULL) {
            plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
            return;
        }
        work = own;
    }
    double *dval = own != NULL ? own + wd : NULL;
    if (dval == NULL) {
        dval = plasma_dev_alloc(sizeof(double));
        if (dval == NULL) {
            plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
            return;
        }
    }
    int err;
    if (empty) {
        err = plasma_kern_fill_int((int *)dval, 0, 2, ctx->stream);
    }
    else {
        int slot;
        const double bytes = (double)plasma_desc_storage_bytes(A);
        plasma_prof_begin(ctx->stream, PlasmaProfMixed, 0, bytes, &slot);
        err = plasma_kern_norm(norm, uplo, A, work, dval, st, ctx->stream);
        plasma_prof_end(ctx->stream, slot);
    }
    if (err)
        plasma_request_fail(sequence, request, err);
    else
        deliver(ctx, value, dval, 1, vprec, sequence, request);
    if (own != NULL)
        plasma_dev_free(own);
    else
        plasma_dev_free(dval);
}

static void omp_amax(plasma_enum_t colrow, plasma_desc_t A, void *values,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (colrow != PlasmaColumnwise && colrow != PlasmaRowwise) {
        plasma_error("illegal value of colrow");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (plasma_desc_check(A) != PlasmaSuccess) {
        plasma_error("invalid descriptor A");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (sequence == NULL) {
        plasma_error("NULL sequence");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (request == NULL) {
        plasma_error("NULL request");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (values == NULL) {
        plasma_error("NULL values");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!tile_aligned(A)) {
        plasma_error("views must start on a tile boundary");
        plasma_request_fail(sequence, request, PlasmaErrorNotSupported);
        return;
    }
    /* quick return (reference dzamax.c:139) */
    if (imin(A.m, A.n) == 0 || sequence->status != PlasmaSuccess)
        return;
    int *st = plasma_sequence_dstatus(sequence, request);
    const size_t count = colrow == PlasmaColumnwise ? (size_t)A.n
                                                    : (size_t)A.mt * A.mb;
    double *dv = plasma_dev_alloc(count * sizeof(double));
    if (dv == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    int slot;
    plasma_prof_begin(ctx->stream, PlasmaProfMixed, 0,
                      (double)plasma_desc_storage_bytes(A), &slot);
    int err = plasma_kern_amax(colrow, A, dv, st, ctx->stream);
    plasma_prof_end(ctx->stream, slot);
    if (err)
        plasma_request_fail(sequence, request, err);
    else
        deliver(ctx, values, dv,
                colrow == PlasmaColumnwise ? (size_t)A.n : (size_t)A.m,
                A.precision, sequence, request);
    plasma_dev_free(dv);
}

/* ---- synchronous drivers ------------------------------------------------ */

static double sync_norm(plasma_enum_t precision, plasma_enum_t norm,
                        plasma_enum_t uplo, int m, int n, void *pA, int lda)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    const int sym = uplo != PlasmaGeneral;
    if (!valid_norm(norm)) {
        plasma_error("illegal value of norm");
        return -1;
    }
    if (sym) {
        /* plasma_dlansy(norm, uplo, n, pA, lda): codes -1 .. -5 */
        if (uplo != PlasmaUpper && uplo != PlasmaLower) {
            plasma_error("illegal value of uplo");
            return -2;
        }
        if (n < 0) {
            plasma_error("illegal value of n");
            return -3;
        }
        if (lda < imax(1, n)) {
            plasma_error("illegal value of lda");
            return -5;
        }
        if (n == 0)
            return 0.0;
    }
    else {
        /* plasma_dlange(norm, m, n, pA, lda): codes -1 .. -5 */
        if (m < 0) {
            plasma_error("illegal value of m");
            return -2;
        }
        if (n < 0) {
            plasma_error("illegal value of n");
            return -3;
        }
        if (lda < imax(1, m)) {
            plasma_error("illegal value of lda");
            return -5;
        }
        if (imin(m, n) == 0)
            return 0.0;
    }
    const int nb = ctx->nb;
    plasma_desc_t A;
    int retval = sym
        ? plasma_desc_triangular_create(precision, uplo, nb, nb, n, n, 0, 0,
                                        n, n, &A)
        : plasma_desc_general_create(precision, nb, nb, m, n, 0, 0, m, n,
                                     &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_*_create() failed");
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    double value = 0.0;
    plasma_pxlapack2desc(sym ? uplo : PlasmaGeneral, pA, lda, A, &sequence,
                         &request);
    /* computed in double for both precisions, delivered to the host */
    omp_norm(norm, uplo, A, NULL, &value, PlasmaRealDouble, &sequence,
             &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_desc_destroy(&A);
    return status == PlasmaSuccess ? value : (double)status;
}

double plasma_dlansy(plasma_enum_t norm, plasma_enum_t uplo, int n,
                     double *pA, int lda)
{
    return sync_norm(PlasmaRealDouble, norm, uplo, n, n, pA, lda);
}

float plasma_slansy(plasma_enum_t norm, plasma_enum_t uplo, int n, float *pA,
                    int lda)
{
    return (float)sync_norm(PlasmaRealFloat, norm, uplo, n, n, pA, lda);
}

double plasma_dlange(plasma_enum_t norm, int m, int n, double *pA, int lda)
{
    return sync_norm(PlasmaRealDouble, norm, PlasmaGeneral, m, n, pA, lda);
}

float plasma_slange(plasma_enum_t norm, int m, int n, float *pA, int lda)
{
    return (float)sync_norm(PlasmaRealFloat, norm, PlasmaGeneral, m, n, pA,
                            lda);
}

static int sync_amax(plasma_enum_t precision, plasma_enum_t colrow, int m,
                     int n, void *pA, int lda, void *values)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    if (colrow != PlasmaColumnwise && colrow != PlasmaRowwise) {
        plasma_error("illegal value of colrow");
        return -1;
    }
    if (m < 0) {
        plasma_error("illegal value of m");
        return -2;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -3;
    }
    if (lda < imax(1, m)) {
        plasma_error("illegal value of lda");
        return -5;
    }
    if (values == NULL) {
        plasma_error("NULL values");
        return -6;
    }
    if (imin(m, n) == 0)
        return PlasmaSuccess;
    const int nb = ctx->nb;
    plasma_desc_t A;
    int retval = plasma_desc_general_create(precision, nb, nb, m, n, 0, 0, m,
                                            n, &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    plasma_pxlapack2desc(PlasmaGeneral, pA, lda, A, &sequence, &request);
    omp_amax(colrow, A, values, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_desc_destroy(&A);
    return status;
}

int plasma_damax(plasma_enum_t colrow, int m, int n, double *pA, int lda,
                 double *values)
{
    return sync_amax(PlasmaRealDouble, colrow, m, n, pA, lda, values);
}

int plasma_samax(plasma_enum_t colrow, int m, int n, float *pA, int lda,
                 float *values)
{
    return sync_amax(PlasmaRealFloat, colrow, m, n, pA, lda, values);
}

/* ---- asynchronous tile forms ------------------------------------------- */

void plasma_omp_dlansy(plasma_enum_t norm, plasma_enum_t uplo,
                       plasma_desc_t A, double *work, double *value,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    omp_norm(norm, uplo, A, work, value, PlasmaRealDouble, sequence, request);
}

void plasma_omp_slansy(plasma_enum_t norm, plasma_enum_t uplo,
                       plasma_desc_t A, double *work, float *value,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    omp_norm(norm, uplo, A, work, value, PlasmaRealFloat, sequence, request);
}

void plasma_omp_dlange(plasma_enum_t norm, plasma_desc_t A, double *work,
                       double *value, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    omp_norm(norm, PlasmaGeneral, A, work, value, PlasmaRealDouble, sequence,
             request);
}

void plasma_omp_slange(plasma_enum_t norm, plasma_desc_t A, double *work,
                       float *value, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    omp_norm(norm, PlasmaGeneral, A, work, value, PlasmaRealFloat, sequence,
             request);
}

void plasma_omp_damax(plasma_enum_t colrow, plasma_desc_t A, double *work,
                      double *values, plasma_sequence_t *sequence,
                      plasma_request_t *request)
{
    (void)work;
    omp_amax(colrow, A, values, sequence, request);
}

void plasma_omp_samax(plasma_enum_t colrow, plasma_desc_t A, double *work,
                      float *values, plasma_sequence_t *sequence,
                      plasma_request_t *request)
{
    (void)work;
    omp_amax(colrow, A, values, sequence, request);
}
