/*
 * potri.c -- the inverse of an SPD matrix from its Cholesky factor:
 * plasma_{d,s}potri (A holds the factor) and plasma_{d,s}poinv (A holds
 * the matrix; factor first), with their tile forms plasma_omp_*.
 *
 * Signatures, argument checks, descriptors and return codes follow the
 * reference (compute/zpotri.c:65-186: a general tile descriptor; compute/
 * zpoinv.c:67: a triangular one; info = the order of the first zero on the
 * factor's diagonal, or potrf's info).  The reference forms L^{-1} in
 * place by the tile algorithm of compute/pztrtri.c and then L^{-T} L^{-1}
 * by compute/pzlauum.c.  Here both products run as whole-matrix MFMA
 * launches:
 *   1. the diagonal-tile inverses W_kk (from the factorization, or formed
 *      for a given factor after a check for a zero diagonal);
 *   2. W = L^{-1} by the forward sweep of the identity (the wide-right-
 *      hand-side sweep: each step one 128 x 128-block MFMA product);
 *   3. Q = W^T W, the triangular Gram product (tiles I >= J of
 *      sum_{K >= I} W(K,I)^T W(K,J), mirrored into a symmetric Q);
 *   4. the uplo triangle of Q into A.
 * W and Q are n x n workspaces from the library's pool.
 */
#include "internal.h"

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err)
        plasma_request_fail(sequence, request, err);
    return err;
}

static void omp_inverse(plasma_enum_t precision, plasma_enum_t uplo,
                        plasma_desc_t A, int factor,
                        plasma_sequence_t *sequence,
                        plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (plasma_desc_check(A) != PlasmaSuccess) {
        plasma_error("invalid A");
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
    if (A.precision != precision || A.mb != A.nb || A.m != A.n) {
        plasma_error("A must be square with square tiles of the routine's "
                     "precision");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    /* quick return (reference zpotri.c: A.n == 0) */
    if (A.n == 0 || sequence->status != PlasmaSuccess)
        return;

    hipStream_t s = ctx->stream;
    int *st = plasma_sequence_dstatus(sequence, request);
    const int nb = A.nb, n = A.n;
    plasma_desc_t W, Q;
    if (plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n, &W) !=
        PlasmaSuccess) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    if (plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n, &Q) !=
        PlasmaSuccess) {
        plasma_desc_destroy(&W);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    void *inv = plasma_dev_alloc(plasma_inv_bytes(A));
    if (inv == NULL) {
        plasma_desc_destroy(&W);
        plasma_desc_destroy(&Q);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }

    /* 1. W_kk: from the factorization, or for the given factor */
    if (factor)
        plasma_pxpotrf(uplo, A, inv, sequence, request);
    else if (!issue(plasma_kern_diag_zero(uplo, A, st, s), sequence,
                    request))
        issue(plasma_kern_trtri_diag(uplo, A, inv, st, s), sequence,
              request);
    /* 2. W = L^{-1}: the forward sweep (L = U^T for upper storage) of I */
    if (sequence->status == PlasmaSuccess &&
        !issue(plasma_kern_identity(W, st, s), sequence, request))
        plasma_pxtrsm_left(uplo, uplo == PlasmaUpper ? PlasmaConjTrans
                                                     : PlasmaNoTrans,
                           A, W, inv, sequence, request);
    /* 3. Q = W^T W (W is lower triangular, held in a general descriptor);
     * 4. A's triangle from Q */
    if (sequence->status == PlasmaSuccess &&
        !issue(plasma_kern_ltl_lower(PlasmaLower, W, W, Q, st, s), sequence,
               request))
        issue(plasma_kern_copy_tri(uplo, Q, A, st, s), sequence, request);
    /* the pool hands the blocks out again in stream order */
    plasma_dev_free(inv);
    plasma_desc_destroy(&W);
    plasma_desc_destroy(&Q);
}

static int sync_inverse(plasma_enum_t precision, plasma_enum_t uplo, int n,
                        void *pA, int lda, int factor)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
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
    if (lda < imax(1, n)) {
        plasma_error("illegal value of lda");
        return -4;
    }
    if (n == 0)
        return PlasmaSuccess;
    /* the factorization's tile size (plasma_tune_potrf) for poinv; the
     * context's for potri, whose input is already factored */
    if (factor && ctx->tuning)
        plasma_tune_potrf(ctx, precision, n);
    const int nb = ctx->nb;
    plasma_desc_t A;
    int retval = factor
        ? plasma_desc_triangular_create(precision, uplo, nb, nb, n, n, 0, 0,
                                        n, n, &A)
        : plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n,
                                     &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_*_create() failed");
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    const plasma_enum_t region = factor ? uplo : PlasmaGeneral;
    plasma_pxlapack2desc(region, pA, lda, A, &sequence, &request);
    omp_inverse(precision, uplo, A, factor, &sequence, &request);
    plasma_pxdesc2lapack(region, A, pA, lda, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_desc_destroy(&A);
    return status;
}

int plasma_dpotri(plasma_enum_t uplo, int n, double *pA, int lda)
{
    return sync_inverse(PlasmaRealDouble, uplo, n, pA, lda, 0);
}

int plasma_spotri(plasma_enum_t uplo, int n, float *pA, int lda)
{
    return sync_inverse(PlasmaRealFloat, uplo, n, pA, lda, 0);
}

int plasma_dpoinv(plasma_enum_t uplo, int n, double *pA, int lda)
{
    return sync_inverse(PlasmaRealDouble, uplo, n, pA, lda, 1);
}

int plasma_spoinv(plasma_enum_t uplo, int n, float *pA, int lda)
{
    return sync_inverse(PlasmaRealFloat, uplo, n, pA, lda, 1);
}

void plasma_omp_dpotri(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_inverse(PlasmaRealDouble, uplo, A, 0, sequence, request);
}

void plasma_omp_spotri(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_inverse(PlasmaRealFloat, uplo, A, 0, sequence, request);
}

void plasma_omp_dpoinv(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_inverse(PlasmaRealDouble, uplo, A, 1, sequence, request);
}

void plasma_omp_spoinv(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_inverse(PlasmaRealFloat, uplo, A, 1, sequence, request);
}
