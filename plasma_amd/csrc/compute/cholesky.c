/*
 * cholesky.c -- the LAPACK-layout and tile-async drivers of the Cholesky
 * solve path: plasma_{d,s,z,c}potrf, plasma_{d,s,z,c}potrs,
 * plasma_{d,s,z,c}posv and their plasma_omp_* forms (complex: Hermitian,
 * A = L L^H or U^H U).
 *
 * Argument checks, return codes and quick returns follow the reference
 * drivers (compute/zpotrf.c:74-150 and :206-246, compute/zpotrs.c:77-172
 * and :226-283, compute/zposv.c:90-186 and :244-301), including their
 * differences: potrf and potrs terminate the process when the library is
 * not initialised (plasma_fatal_error) while posv returns
 * PlasmaErrorNotInitialized; plasma_omp_posv reports an illegal uplo
 * without failing the sequence.
 *
 * Internal storage: A is held in a triangular tile descriptor for potrf,
 * potrs and posv (the reference uses a general one for potrs and copies the
 * unused triangle too; only the uplo triangle is ever read).
 */
#include "internal.h"

int plasma_arg_stage_in(lapack_arg_t *a, void *p, int ld, int m, int n,
                        size_t es, int copy_in)
{
    plasma_context_t *ctx = plasma_context_self();
    a->user = p;
    a->user_ld = ld;
    a->m = m;
    a->n = n;
    a->es = es;
    if (plasma_is_device_ptr(p)) {
        a->dev = p;
        a->ld = ld;
        a->staged = 0;
        return PlasmaSuccess;
    }
    a->dev = plasma_dev_alloc((size_t)m * n * es);
    if (a->dev == NULL)
        return PlasmaErrorOutOfMemory;
    a->ld = m;
    a->staged = 1;
    if (copy_in &&
        plasma_hip_check(hipMemcpy2DAsync(a->dev, (size_t)m * es, p,
                                          (size_t)ld * es, (size_t)m * es, n,
                                          hipMemcpyHostToDevice, ctx->stream),
                         "hipMemcpy2D")) {
        plasma_dev_free(a->dev);
        return PlasmaErrorInternal;
    }
    return PlasmaSuccess;
}

void plasma_arg_stage_out(lapack_arg_t *a, int copy_back)
{
    if (!a->staged)
        return;
    if (copy_back)
        plasma_hip_check(hipMemcpy2D(a->user, (size_t)a->user_ld * a->es,
                                     a->dev, (size_t)a->m * a->es,
                                     (size_t)a->m * a->es, a->n,
                                     hipMemcpyDeviceToHost), "hipMemcpy2D");
    plasma_dev_free(a->dev);
}

/* Tile size for a factorization of order n (reference: plasma_tune_potrf,
 * control/tuning.c:367-374): with tuning enabled the autotuned table, when
 * present, sets the context's nb (and schedule) first. */
static int tuned_nb(plasma_context_t *ctx, plasma_enum_t precision, int n)
{
    plasma_tune_potrf(ctx, precision, n);
    return ctx->nb;
}

/***************************************************************************/
void plasma_omp_xpotrf(plasma_enum_t precision, plasma_enum_t uplo,
                       plasma_desc_t A, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    if (plasma_context_self() == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        return;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (plasma_desc_check(A) != PlasmaSuccess) {
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        plasma_error("invalid A");
        return;
    }
    if (sequence == NULL) {
        plasma_fatal_error("NULL sequence");
        return;
    }
    if (request == NULL) {
        plasma_fatal_error("NULL request");
        return;
    }
    if (A.m == 0)
        return;
    if (A.precision != precision || A.mb != A.nb) {
        plasma_error("A must hold square tiles of the routine's precision");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    plasma_pxpotrf(uplo, A, NULL, sequence, request);
}

static int check_pair(plasma_enum_t precision, plasma_desc_t A,
                      plasma_desc_t B, plasma_sequence_t *sequence,
                      plasma_request_t *request)
{
    if (plasma_desc_check(A) != PlasmaSuccess) {
        plasma_error("invalid A");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (plasma_desc_check(B) != PlasmaSuccess) {
        plasma_error("invalid B");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    if (sequence == NULL) {
        plasma_fatal_error("NULL sequence");
        return 0;
    }
    if (request == NULL) {
        plasma_fatal_error("NULL request");
        return 0;
    }
    if (A.precision != precision || B.precision != precision ||
        A.mb != A.nb || B.mb != A.mb || A.m != A.n || B.m != A.m) {
        plasma_error("A and B must share the tile size and precision");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    return 1;
}

void plasma_omp_xpotrs(plasma_enum_t precision, plasma_enum_t uplo,
                       plasma_desc_t A, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (plasma_context_self() == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        return;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!check_pair(precision, A, B, sequence, request))
        return;
    if (A.n == 0 || B.n == 0)
        return;
    plasma_pxpotrs(uplo, A, B, NULL, sequence, request);
}

void plasma_omp_xposv(plasma_enum_t precision, plasma_enum_t uplo,
                      plasma_desc_t A, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!check_pair(precision, A, B, sequence, request))
        return;
    if (A.n == 0 || B.n == 0)
        return;
    /* the factorization keeps the inverses of its diagonal tiles for the
     * two sweeps */
    void *inv = plasma_dev_alloc(plasma_inv_bytes(A));
    if (inv == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    plasma_pxpotrf(uplo, A, inv, sequence, request);
    plasma_pxpotrs(uplo, A, B, inv, sequence, request);
    plasma_dev_free(inv);
}

/***************************************************************************/
int plasma_xpotrf(plasma_enum_t precision, plasma_enum_t uplo, int n,
                  void *pA, int lda)
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

    int nb = ctx->tuning ? tuned_nb(ctx, precision, n) : ctx->nb;
    size_t es = plasma_element_size(precision);
    plasma_desc_t A;
    int retval = plasma_desc_triangular_create(precision, uplo, nb, nb, n, n,
                                               0, 0, n, n, &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_triangular_create() failed");
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);

    lapack_arg_t a;
    retval = plasma_arg_stage_in(&a, pA, lda, n, n, es, 1);
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        return retval;
    }
    plasma_pxlapack2desc(uplo, a.dev, a.ld, A, &sequence, &request);
    plasma_omp_xpotrf(precision, uplo, A, &sequence, &request);
    plasma_pxdesc2lapack(uplo, A, a.dev, a.ld, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_arg_stage_out(&a, status == PlasmaSuccess);
    plasma_desc_destroy(&A);
    return status;
}

static int solve_args(plasma_enum_t uplo, int n, int nrhs, int lda, int ldb)
{
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
    if (ldb < imax(1, n)) {
        plasma_error("illegal value of ldb");
        return -7;
    }
    return 0;
}

/* factor: 1 = posv (factor then solve), 0 = potrs (solve only) */
static int solve_driver(plasma_enum_t precision, plasma_enum_t uplo, int n,
                        int nrhs, void *pA, int lda, void *pB, int ldb,
                        int factor)
{
    plasma_context_t *ctx = plasma_context_self();
    int nb = (factor && ctx->tuning) ? tuned_nb(ctx, precision, n) : ctx->nb;
    size_t es = plasma_element_size(precision);
    plasma_desc_t A, B;
    int retval = plasma_desc_triangular_create(precision, uplo, nb, nb, n, n,
                                               0, 0, n, n, &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_triangular_create() failed");
        return retval;
    }
    retval = plasma_desc_general_create(precision, nb, nb, n, nrhs, 0, 0, n,
                                        nrhs, &B);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        plasma_desc_destroy(&A);
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);

    lapack_arg_t a, b;
    retval = plasma_arg_stage_in(&a, pA, lda, n, n, es, 1);
    if (retval == PlasmaSuccess) {
        retval = plasma_arg_stage_in(&b, pB, ldb, n, nrhs, es, 1);
        if (retval != PlasmaSuccess)
            plasma_arg_stage_out(&a, 0);
    }
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        plasma_desc_destroy(&B);
        return retval;
    }

    plasma_pxlapack2desc(uplo, a.dev, a.ld, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, b.dev, b.ld, B, &sequence, &request);
    if (factor) {
        plasma_omp_xposv(precision, uplo, A, B, &sequence, &request);
        plasma_pxdesc2lapack(uplo, A, a.dev, a.ld, &sequence, &request);
    }
    else {
        plasma_omp_xpotrs(precision, uplo, A, B, &sequence, &request);
    }
    plasma_pxdesc2lapack(PlasmaGeneral, B, b.dev, b.ld, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_arg_stage_out(&a, factor && status == PlasmaSuccess);
    plasma_arg_stage_out(&b, status == PlasmaSuccess);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&B);
    return status;
}

int plasma_xpotrs(plasma_enum_t precision, plasma_enum_t uplo, int n,
                  int nrhs, void *pA, int lda, void *pB, int ldb)
{
    if (plasma_context_self() == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    int err = solve_args(uplo, n, nrhs, lda, ldb);
    if (err)
        return err;
    if (imin(n, nrhs) == 0)
        return PlasmaSuccess;
    return solve_driver(precision, uplo, n, nrhs, pA, lda, pB, ldb, 0);
}

int plasma_xposv(plasma_enum_t precision, plasma_enum_t uplo, int n,
                 int nrhs, void *pA, int lda, void *pB, int ldb)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    int err = solve_args(uplo, n, nrhs, lda, ldb);
    if (err)
        return err;
    if (imin(n, nrhs) == 0)
        return PlasmaSuccess;
    return solve_driver(precision, uplo, n, nrhs, pA, lda, pB, ldb, 1);
}

/***************************************************************************
 * Precision entry points.
 */
#define DEFINE_CHOL(P, TYPE, PREC)                                            \
int plasma_##P##potrf(plasma_enum_t uplo, int n, TYPE *pA, int lda)           \
{                                                                             \
    return plasma_xpotrf(PREC, uplo, n, pA, lda);                             \
}                                                                             \
int plasma_##P##potrs(plasma_enum_t uplo, int n, int nrhs, TYPE *pA, int lda, \
                      TYPE *pB, int ldb)                                      \
{                                                                             \
    return plasma_xpotrs(PREC, uplo, n, nrhs, pA, lda, pB, ldb);              \
}                                                                             \
int plasma_##P##posv(plasma_enum_t uplo, int n, int nrhs, TYPE *pA, int lda,  \
                     TYPE *pB, int ldb)                                       \
{                                                                             \
    return plasma_xposv(PREC, uplo, n, nrhs, pA, lda, pB, ldb);               \
}                                                                             \
void plasma_omp_##P##potrf(plasma_enum_t uplo, plasma_desc_t A,               \
                           plasma_sequence_t *sequence,                       \
                           plasma_request_t *request)                         \
{                                                                             \
    plasma_omp_xpotrf(PREC, uplo, A, sequence, request);                      \
}                                                                             \
void plasma_omp_##P##potrs(plasma_enum_t uplo, plasma_desc_t A,               \
                           plasma_desc_t B, plasma_sequence_t *sequence,      \
                           plasma_request_t *request)                         \
{                                                                             \
    plasma_omp_xpotrs(PREC, uplo, A, B, sequence, request);                   \
}                                                                             \
void plasma_omp_##P##posv(plasma_enum_t uplo, plasma_desc_t A,                \
                          plasma_desc_t B, plasma_sequence_t *sequence,       \
                          plasma_request_t *request)                          \
{                                                                             \
    plasma_omp_xposv(PREC, uplo, A, B, sequence, request);                    \
}

DEFINE_CHOL(d, double, PlasmaRealDouble)
DEFINE_CHOL(s, float, PlasmaRealFloat)
DEFINE_CHOL(z, plasma_complex64_t, PlasmaComplexDouble)
DEFINE_CHOL(c, plasma_complex32_t, PlasmaComplexFloat)
