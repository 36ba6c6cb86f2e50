/*
 * adjoint.c -- reverse-mode derivatives of the Cholesky solve path:
 * plasma_{d,s}potrf_adjoint and plasma_{d,s}potrs_adjoint.  The reference
 * has no derivatives; these serve callers that differentiate through a
 * factorization (Gaussian-process log-likelihoods, implicit layers).
 *
 * Gradients are symmetric: Abar is the full n x n matrix with
 * df = sum_ij Abar(i,j) dA(i,j) for every symmetric perturbation dA.
 *
 * Cholesky adjoint.  With A = L L^T and Lbar the adjoint of the factor,
 *     Abar = 1/2 L^{-T} Q L^{-1},
 * Q the symmetric matrix whose lower triangle is tril(L^T Lbar) (the
 * symmetric form of I. Murray, "Differentiation of the Cholesky
 * decomposition", 2016, eq. 10: Phi(L^T Lbar) + Phi(L^T Lbar)^T = Q, Phi
 * taking the lower triangle with the diagonal halved).  On the device:
 *     Q  = ltl_lower(L, Lbar)           triangular MFMA product, mirrored
 *     G  = L^{-T} Q                     backward sweep, n right-hand sides
 *     H  = L^{-T} G^T  (= L^{-T} Q L^{-1}, symmetric): lower triangle only,
 *                                       a third of a full sweep
 *     Abar = H / 2, mirrored            exact symmetry
 * The sweeps are the potrs kernels with the diagonal-tile inverses (block
 * GEMMs for this many right-hand sides).
 *
 * potrs adjoint.  X = A^{-1} B with adjoint Xbar: Bbar = A^{-1} Xbar and
 * Abar -= 1/2 (Bbar X^T + X Bbar^T).
 */
#include "internal.h"

static int adj_args_potrf(plasma_enum_t uplo, int n, int lda, int ldabar)
{
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
    if (ldabar < imax(1, n)) {
        plasma_error("illegal value of ldabar");
        return -6;
    }
    return 0;
}

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err)
        plasma_request_fail(sequence, request, err);
    return err;
}

int plasma_xpotrf_adjoint(plasma_enum_t precision, plasma_enum_t uplo,
                          int n, const void *pA, int lda, void *pAbar,
                          int ldabar)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    int err = adj_args_potrf(uplo, n, lda, ldabar);
    if (err)
        return err;
    if (n == 0)
        return PlasmaSuccess;
    if (ctx->tuning)
        plasma_tune_potrf(ctx, precision, n);
    /* the sweeps with n right-hand sides dominate: their block GEMMs want
     * tiles of at least 256 once n is large (n = 16384, FP64: 250 ms at nb
     * 128, 218 ms at 256) */
    const int nb = (n >= 8192) ? imax(ctx->nb, 256) : ctx->nb;
    const size_t es = plasma_element_size(precision);

    plasma_desc_t A, Lb, Q;
    int retval = plasma_desc_triangular_create(precision, uplo, nb, nb, n, n,
                                               0, 0, n, n, &A);
    if (retval != PlasmaSuccess)
        return retval;
    retval = plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n,
                                        &Lb);
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        return retval;
    }
    retval = plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n,
                                        &Q);
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        plasma_desc_destroy(&Lb);
        return retval;
    }
    void *inv = plasma_dev_alloc(plasma_inv_bytes(A));
    lapack_arg_t a, ab;
    retval = inv ? plasma_arg_stage_in(&a, (void *)pA, lda, n, n, es, 1)
                 : PlasmaErrorOutOfMemory;
    if (retval == PlasmaSuccess) {
        retval = plasma_arg_stage_in(&ab, pAbar, ldabar, n, n, es, 1);
        if (retval != PlasmaSuccess)
            plasma_arg_stage_out(&a, 0);
    }
    if (retval != PlasmaSuccess) {
        plasma_dev_free(inv);
        plasma_desc_destroy(&A);
        plasma_desc_destroy(&Lb);
        plasma_desc_destroy(&Q);
        return retval;
    }

    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    hipStream_t s = ctx->stream;
    /* the solve with L^T: backward sweep */
    const plasma_enum_t back = (uplo == PlasmaLower) ? PlasmaConjTrans
                                                     : PlasmaNoTrans;

    plasma_pxlapack2desc(uplo, a.dev, a.ld, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, ab.dev, ab.ld, Lb, &sequence,
                         &request);
    int *st = plasma_sequence_dstatus(&sequence, &request);
    if (sequence.status == PlasmaSuccess && st != NULL) {
        int slot;
        const double dn = n;
        plasma_prof_begin(s, PlasmaProfAdjoint, dn * dn * dn / 3.0, 0,
                          &slot);
        int e = issue(plasma_kern_trtri_diag(uplo, A, inv, st, s),
                      &sequence, &request);
        if (!e)
            e = issue(plasma_kern_ltl_lower(uplo, A, Lb, Q, st, s),
                      &sequence, &request);
        plasma_prof_end(s, slot);
        if (!e)
            plasma_pxtrsm_left(uplo, back, A, Q, inv, &sequence, &request);
        if (!e && sequence.status == PlasmaSuccess) {
            plasma_prof_begin(s, PlasmaProfAdjoint, 0,
                              2.0 * dn * dn * es, &slot);
            e = issue(plasma_kern_tile_axpby_t(0.0, Q, 1.0, Q, Lb, 0, st,
                                               s),
                      &sequence, &request);
            plasma_prof_end(s, slot);
        }
        if (!e)
            plasma_pxtrsm_left_lower(uplo, A, Lb, inv, &sequence, &request);
        if (!e && sequence.status == PlasmaSuccess) {
            plasma_prof_begin(s, PlasmaProfAdjoint, 0,
                              2.0 * dn * dn * es, &slot);
            issue(plasma_kern_tile_axpby_t(0.5, Lb, 0.5, Lb, Q, 1, st, s),
                  &sequence, &request);
            plasma_prof_end(s, slot);
        }
    }
    plasma_pxdesc2lapack(PlasmaGeneral, Q, ab.dev, ab.ld, &sequence,
                         &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_arg_stage_out(&a, 0);
    plasma_arg_stage_out(&ab, status == PlasmaSuccess);
    plasma_dev_free(inv);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&Lb);
    plasma_desc_destroy(&Q);
    return status;
}

int plasma_xpotrs_adjoint(plasma_enum_t precision, plasma_enum_t uplo,
                          int n, int nrhs, const void *pA, int lda,
                          const void *pX, int ldx, void *pXbar, int ldxbar,
                          void *pAbar, int ldabar)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
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
    if (nrhs < 0) {
        plasma_error("illegal value of nrhs");
        return -3;
    }
    if (lda < imax(1, n)) {
        plasma_error("illegal value of lda");
        return -5;
    }
    if (ldx < imax(1, n)) {
        plasma_error("illegal value of ldx");
        return -7;
    }
    if (ldxbar < imax(1, n)) {
        plasma_error("illegal value of ldxbar");
        return -9;
    }
    if (ldabar < imax(1, n)) {
        plasma_error("illegal value of ldabar");
        return -11;
    }
    if (imin(n, nrhs) == 0)
        return PlasmaSuccess;
    const int nb = ctx->nb;
    const size_t es = plasma_element_size(precision);

    plasma_desc_t A, B;
    int retval = plasma_desc_triangular_create(precision, uplo, nb, nb, n, n,
                                               0, 0, n, n, &A);
    if (retval != PlasmaSuccess)
        return retval;
    retval = plasma_desc_general_create(precision, nb, nb, n, nrhs, 0, 0, n,
                                        nrhs, &B);
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        return retval;
    }
    lapack_arg_t a, x, xb, ab;
    int staged = 0;
    retval = plasma_arg_stage_in(&a, (void *)pA, lda, n, n, es, 1);
    if (retval == PlasmaSuccess) {
        staged = 1;
        retval = plasma_arg_stage_in(&x, (void *)pX, ldx, n, nrhs, es, 1);
    }
    if (retval == PlasmaSuccess) {
        staged = 2;
        retval = plasma_arg_stage_in(&xb, pXbar, ldxbar, n, nrhs, es, 1);
    }
    if (retval == PlasmaSuccess) {
        staged = 3;
        retval = plasma_arg_stage_in(&ab, pAbar, ldabar, n, n, es, 1);
    }
    if (retval != PlasmaSuccess) {
        if (staged > 2) plasma_arg_stage_out(&xb, 0);
        if (staged > 1) plasma_arg_stage_out(&x, 0);
        if (staged > 0) plasma_arg_stage_out(&a, 0);
        plasma_desc_destroy(&A);
        plasma_desc_destroy(&B);
        return retval;
    }

    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    plasma_pxlapack2desc(uplo, a.dev, a.ld, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, xb.dev, xb.ld, B, &sequence,
                         &request);
    plasma_pxpotrs(uplo, A, B, NULL, &sequence, &request);
    plasma_pxdesc2lapack(PlasmaGeneral, B, xb.dev, xb.ld, &sequence,
                         &request);
    if (sequence.status == PlasmaSuccess) {
        int slot;
        plasma_prof_begin(ctx->stream, PlasmaProfAdjoint,
                          4.0 * n * (double)n * nrhs, 0, &slot);
        issue(plasma_kern_syr2k_lapack(precision, n, nrhs, 0.5, xb.dev,
                                       xb.ld, x.dev, x.ld, ab.dev, ab.ld,
                                       ctx->stream),
              &sequence, &request);
        plasma_prof_end(ctx->stream, slot);
    }
    int status = plasma_sequence_sync(&sequence);
    plasma_arg_stage_out(&a, 0);
    plasma_arg_stage_out(&x, 0);
    plasma_arg_stage_out(&xb, status == PlasmaSuccess);
    plasma_arg_stage_out(&ab, status == PlasmaSuccess);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&B);
    return status;
}

#define DEFINE_ADJ(P, TYPE, PREC)                                             \
int plasma_##P##potrf_adjoint(plasma_enum_t uplo, int n, const TYPE *pA,      \
                              int lda, TYPE *pAbar, int ldabar)               \
{                                                                             \
    return plasma_xpotrf_adjoint(PREC, uplo, n, pA, lda, pAbar, ldabar);      \
}                                                                             \
int plasma_##P##potrs_adjoint(plasma_enum_t uplo, int n, int nrhs,            \
                              const TYPE *pA, int lda, const TYPE *pX,        \
                              int ldx, TYPE *pXbar, int ldxbar, TYPE *pAbar,  \
                              int ldabar)                                     \
{                                                                             \
    return plasma_xpotrs_adjoint(PREC, uplo, n, nrhs, pA, lda, pX, ldx,       \
                                 pXbar, ldxbar, pAbar, ldabar);               \
}

DEFINE_ADJ(d, double, PlasmaRealDouble)
DEFINE_ADJ(s, float, PlasmaRealFloat)
