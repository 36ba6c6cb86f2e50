/*
 * band.c -- Cholesky factorization and solve of a symmetric positive
 * definite band matrix: plasma_{d,s}pbtrf, plasma_{d,s}pbtrs,
 * plasma_{d,s}pbsv, their plasma_omp_* tile forms and the band layout
 * translations plasma_omp_{d,s}{pb2desc,desc2pb}.
 *
 * Argument checks, return codes and quick returns follow the reference
 * drivers (compute/zpbtrf.c:77-150 and :200-242, compute/zpbtrs.c:81-180
 * and :236-295, compute/zpbsv.c:92-190, compute/zpb2desc.c:27-68,
 * compute/zdesc2pb.c:27-68), including pbtrs's codes (-4 for kd, -5 for
 * nrhs, -7 for ldab, -10 for ldb).  The LAPACK band array AB holds the kd
 * + 1 diagonals of the stored triangle (column j: A(i,j) at AB(kd+1+i-j, j)
 * for upper, AB(1+i-j, j) for lower, 1-based); it may be host or device
 * memory.  Internally the band lives in a band tile descriptor of
 * nb (1 + ceil(kd / nb)) rows (reference zpbtrf.c:104-110).
 */
#include "internal.h"

/* ---- layout ------------------------------------------------------------ */

static int band_kd(plasma_desc_t A)
{
    return A.uplo == PlasmaUpper ? A.ku : A.kl;
}

static void band_layout(plasma_desc_t A, void *pAB, int ldab, int to_tiles,
                        plasma_sequence_t *sequence,
                        plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess || A.m == 0)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    const size_t es = plasma_element_size(A.precision);
    lapack_arg_t ab;
    int err = plasma_arg_stage_in(&ab, pAB, ldab, ldab, A.n, es, 1);
    if (err != PlasmaSuccess) {
        plasma_request_fail(sequence, request, err);
        return;
    }
    int slot;
    plasma_prof_begin(ctx->stream, PlasmaProfLayout, 0,
                      2.0 * (band_kd(A) + 1.0) * A.n * es, &slot);
    err = plasma_kern_band_layout(A, ab.dev, ab.ld, band_kd(A),
                                  A.uplo == PlasmaUpper, to_tiles, st,
                                  ctx->stream);
    plasma_prof_end(ctx->stream, slot);
    if (err)
        plasma_request_fail(sequence, request, err);
    if (ab.staged) {
        /* towards the host only after the work queued before it */
        if (!to_tiles)
            hipStreamSynchronize(ctx->stream);
        plasma_arg_stage_out(&ab, !to_tiles && sequence->status ==
                                                   PlasmaSuccess);
    }
}

static int layout_args(plasma_enum_t precision, void *pA, int lda,
                       plasma_desc_t A, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (plasma_desc_check(A) != PlasmaSuccess) {
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        plasma_error("invalid A");
        return 0;
    }
    if (sequence == NULL) {
        plasma_fatal_error("NULL sequence");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (request == NULL) {
        plasma_fatal_error("NULL request");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (A.type != PlasmaGeneralBand || A.precision != precision ||
        A.mb != A.nb || (A.uplo != PlasmaUpper && A.uplo != PlasmaLower)) {
        plasma_error("A must be a symmetric band descriptor of the "
                     "routine's precision");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (pA == NULL || lda < band_kd(A) + 1) {
        plasma_error("illegal band array");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    return 1;
}

static void omp_pb2desc(plasma_enum_t precision, void *pA, int lda,
                        plasma_desc_t A, plasma_sequence_t *sequence,
                        plasma_request_t *request)
{
    if (layout_args(precision, pA, lda, A, sequence, request))
        band_layout(A, pA, lda, 1, sequence, request);
}

static void omp_desc2pb(plasma_enum_t precision, plasma_desc_t A, void *pA,
                        int lda, plasma_sequence_t *sequence,
                        plasma_request_t *request)
{
    if (layout_args(precision, pA, lda, A, sequence, request))
        band_layout(A, pA, lda, 0, sequence, request);
}

void plasma_omp_dpb2desc(double *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request)
{
    omp_pb2desc(PlasmaRealDouble, pA, lda, A, sequence, request);
}

void plasma_omp_spb2desc(float *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request)
{
    omp_pb2desc(PlasmaRealFloat, pA, lda, A, sequence, request);
}

void plasma_omp_ddesc2pb(plasma_desc_t A, double *pA, int lda,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request)
{
    omp_desc2pb(PlasmaRealDouble, A, pA, lda, sequence, request);
}

void plasma_omp_sdesc2pb(plasma_desc_t A, float *pA, int lda,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request)
{
    omp_desc2pb(PlasmaRealFloat, A, pA, lda, sequence, request);
}

/* ---- tile forms --------------------------------------------------------- */

static int omp_args(plasma_enum_t precision, plasma_enum_t uplo,
                    plasma_desc_t AB, const plasma_desc_t *B,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (plasma_context_self() == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (plasma_desc_check(AB) != PlasmaSuccess) {
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        plasma_error("invalid A");
        return 0;
    }
    if (B != NULL && plasma_desc_check(*B) != PlasmaSuccess) {
        plasma_error("invalid B");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (sequence == NULL) {
        plasma_fatal_error("NULL sequence");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (request == NULL) {
        plasma_fatal_error("NULL request");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (AB.type != PlasmaGeneralBand || AB.uplo != uplo ||
        AB.precision != precision || AB.mb != AB.nb || AB.m != AB.n ||
        (B != NULL && (B->precision != precision || B->mb != AB.mb ||
                       B->m != AB.m))) {
        plasma_error("descriptors do not match the band solver contract");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    return 1;
}

/* factor and / or solve on descriptors; inv: the tile inverses, NULL to
 * use a private buffer (computed for a solve without the factorization) */
static void omp_band(plasma_enum_t precision, plasma_enum_t uplo,
                     plasma_desc_t AB, const plasma_desc_t *B, int factor,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (!omp_args(precision, uplo, AB, B, sequence, request))
        return;
    if (AB.m == 0 || (B != NULL && B->n == 0 && !factor))
        return;
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    void *inv = plasma_dev_alloc(plasma_inv_bytes(AB));
    if (inv == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    if (factor) {
        plasma_pxpbtrf(uplo, AB, inv, sequence, request);
    }
    else if (B != NULL && B->n > 0) {
        int *st = plasma_sequence_dstatus(sequence, request);
        int err = plasma_kern_trtri_diag(uplo, AB, inv, st, ctx->stream);
        if (err)
            plasma_request_fail(sequence, request, err);
    }
    if (B != NULL && B->n > 0)
        plasma_pxpbtrs(uplo, AB, *B, inv, sequence, request);
    plasma_dev_free(inv);
}

void plasma_omp_dpbtrf(plasma_enum_t uplo, plasma_desc_t AB,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_band(PlasmaRealDouble, uplo, AB, NULL, 1, sequence, request);
}

void plasma_omp_spbtrf(plasma_enum_t uplo, plasma_desc_t AB,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_band(PlasmaRealFloat, uplo, AB, NULL, 1, sequence, request);
}

void plasma_omp_dpbtrs(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_band(PlasmaRealDouble, uplo, AB, &B, 0, sequence, request);
}

void plasma_omp_spbtrs(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_band(PlasmaRealFloat, uplo, AB, &B, 0, sequence, request);
}

void plasma_omp_dpbsv(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_band(PlasmaRealDouble, uplo, AB, &B, 1, sequence, request);
}

void plasma_omp_spbsv(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_band(PlasmaRealFloat, uplo, AB, &B, 1, sequence, request);
}

/* ---- LAPACK-layout drivers ---------------------------------------------- */

/* factor: 1 pbtrf / pbsv, 0 pbtrs; nrhs < 0: no right-hand sides */
static int band_driver(plasma_enum_t precision, plasma_enum_t uplo, int n,
                       int kd, int nrhs, void *pAB, int ldab, void *pB,
                       int ldb, int factor)
{
    plasma_context_t *ctx = plasma_context_self();
    const int nb = ctx->nb;
    const size_t es = plasma_element_size(precision);
    const int lm = nb * (1 + (kd + nb - 1) / nb);
    plasma_desc_t AB, B;
    int retval = plasma_desc_general_band_create(precision, uplo, nb, nb, lm,
                                                 n, 0, 0, n, n, kd, kd, &AB);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_band_create() failed");
        return retval;
    }
    if (nrhs > 0) {
        retval = plasma_desc_general_create(precision, nb, nb, n, nrhs, 0, 0,
                                            n, nrhs, &B);
        if (retval != PlasmaSuccess) {
            plasma_error("plasma_desc_general_create() failed");
            plasma_desc_destroy(&AB);
            return retval;
        }
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);

    lapack_arg_t ab, b;
    retval = plasma_arg_stage_in(&ab, pAB, ldab, ldab, n, es, 1);
    if (retval == PlasmaSuccess && nrhs > 0) {
        retval = plasma_arg_stage_in(&b, pB, ldb, n, nrhs, es, 1);
        if (retval != PlasmaSuccess)
            plasma_arg_stage_out(&ab, 0);
    }
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&AB);
        if (nrhs > 0)
            plasma_desc_destroy(&B);
        return retval;
    }
    band_layout(AB, ab.dev, ab.ld, 1, &sequence, &request);
    if (nrhs > 0)
        plasma_pxlapack2desc(PlasmaGeneral, b.dev, b.ld, B, &sequence,
                             &request);
    omp_band(precision, uplo, AB, nrhs > 0 ? &B : NULL, factor, &sequence,
             &request);
    if (factor)
        band_layout(AB, ab.dev, ab.ld, 0, &sequence, &request);
    if (nrhs > 0)
        plasma_pxdesc2lapack(PlasmaGeneral, B, b.dev, b.ld, &sequence,
                             &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_arg_stage_out(&ab, factor && status == PlasmaSuccess);
    if (nrhs > 0) {
        plasma_arg_stage_out(&b, status == PlasmaSuccess);
        plasma_desc_destroy(&B);
    }
    plasma_desc_destroy(&AB);
    return status;
}

static int xpbtrf(plasma_enum_t precision, plasma_enum_t uplo, int n, int kd,
                  void *pAB, int ldab)
{
    if (plasma_context_self() == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        return -1;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -2;
    }
    if (kd < 0) {
        plasma_error("illegal value of kd");
        return -3;
    }
    if (ldab < kd + 1) {
        plasma_error("illegal value of ldab");
        return -5;
    }
    if (imax(n, 0) == 0)
        return PlasmaSuccess;
    return band_driver(precision, uplo, n, kd, -1, pAB, ldab, NULL, 0, 1);
}

static int xpbtrs(plasma_enum_t precision, plasma_enum_t uplo, int n, int kd,
                  int nrhs, void *pAB, int ldab, void *pB, int ldb)
{
    if (plasma_context_self() == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        return -1;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -2;
    }
    if (kd < 0) {
        plasma_error("illegal value of kd");
        return -4;
    }
    if (nrhs < 0) {
        plasma_error("illegal value of nrhs");
        return -5;
    }
    if (ldab < imax(1, 1 + kd)) {
        plasma_error("illegal value of ldab");
        return -7;
    }
    if (ldb < imax(1, n)) {
        plasma_error("illegal value of ldb");
        return -10;
    }
    if (imax(n, nrhs) == 0 || imin(n, nrhs) == 0)
        return PlasmaSuccess;
    return band_driver(precision, uplo, n, kd, nrhs, pAB, ldab, pB, ldb, 0);
}

static int xpbsv(plasma_enum_t precision, plasma_enum_t uplo, int n, int kd,
                 int nrhs, void *pAB, int ldab, void *pB, int ldb)
{
    if (plasma_context_self() == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        return -1;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -2;
    }
    if (kd < 0) {
        plasma_error("illegal value of kd");
        return -3;
    }
    if (nrhs < 0) {
        plasma_error("illegal value of nrhs");
        return -4;
    }
    if (ldab < kd + 1) {
        plasma_error("illegal value of ldab");
        return -6;
    }
    if (ldb < imax(1, n)) {
        plasma_error("illegal value of ldb");
        return -8;
    }
    if (imin(n, nrhs) == 0)
        return PlasmaSuccess;
    return band_driver(precision, uplo, n, kd, nrhs, pAB, ldab, pB, ldb, 1);
}

int plasma_dpbtrf(plasma_enum_t uplo, int n, int kd, double *pAB, int ldab)
{
    return xpbtrf(PlasmaRealDouble, uplo, n, kd, pAB, ldab);
}

int plasma_spbtrf(plasma_enum_t uplo, int n, int kd, float *pAB, int ldab)
{
    return xpbtrf(PlasmaRealFloat, uplo, n, kd, pAB, ldab);
}

int plasma_dpbtrs(plasma_enum_t uplo, int n, int kd, int nrhs, double *pAB,
                  int ldab, double *pB, int ldb)
{
    return xpbtrs(PlasmaRealDouble, uplo, n, kd, nrhs, pAB, ldab, pB, ldb);
}

int plasma_spbtrs(plasma_enum_t uplo, int n, int kd, int nrhs, float *pAB,
                  int ldab, float *pB, int ldb)
{
    return xpbtrs(PlasmaRealFloat, uplo, n, kd, nrhs, pAB, ldab, pB, ldb);
}

int plasma_dpbsv(plasma_enum_t uplo, int n, int kd, int nrhs, double *pAB,
                 int ldab, double *pB, int ldb)
{
    return xpbsv(PlasmaRealDouble, uplo, n, kd, nrhs, pAB, ldab, pB, ldb);
}

int plasma_spbsv(plasma_enum_t uplo, int n, int kd, int nrhs, float *pAB,
                 int ldab, float *pB, int ldb)
{
    return xpbsv(PlasmaRealFloat, uplo, n, kd, nrhs, pAB, ldab, pB, ldb);
}
