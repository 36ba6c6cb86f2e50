/*
 * dsposv.c -- mixed-precision Cholesky solve with FP64 iterative refinement
 * (plasma_dsposv, plasma_omp_dsposv; complex: plasma_zcposv,
 * plasma_omp_zcposv) and the FP64 <-> FP32 conversions (plasma_dlag2s,
 * plasma_slag2d, plasma_zlag2c, plasma_clag2z and their tile forms).  One
 * body serves both precision pairs: the kernels dispatch on the
 * descriptors' precision (Hermitian residual and norms for complex data).
 *
 * Algorithm and constants follow the reference (compute/zcposv.c:363-539
 * generated as dsposv):
 *   Anorm = ||A||_inf (symmetric, one triangle); eps = 2^-53;
 *   Xs = fp32(B); As = fp32(A); As = chol(As); Xs = As^{-1} Xs (2 sweeps);
 *   X = fp64(Xs); R = B - A X; stop with iter = 0 if every column j has
 *   max|R_j| <= max|X_j| * Anorm * eps * sqrt(n);
 *   up to 30 times: Xs = fp32(R); Xs = As^{-1} Xs; X += fp64(Xs);
 *   R = B - A X; stop with iter = k + 1 on the same test;
 *   otherwise iter = -31 and the FP64 factorization solves the system.
 * An FP64 -> FP32 overflow (status 1) or a non positive definite FP32
 * factor (status = its info) fails the sequence and ends the computation,
 * as in the reference, where the failed sequence makes every later tile
 * routine return (compute/pzpotrf.c:31-32).
 *
 * All of it runs on the GPU; the host reads back one flag and the status
 * word per iteration (the reference's taskwait at zcposv.c:474, :509).
 * Internally A and As are triangular descriptors: only the uplo triangle is
 * ever read (the reference copies the full matrix, zcposv.c:448 "TODO").
 */
#include <math.h>
#include <string.h>

#include "internal.h"

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err != 0)
        plasma_request_fail(sequence, request, err);
    return err;
}

/* Read the convergence flag and the status word; returns 1 when converged,
 * 0 when not, -1 when the sequence failed. */
static int read_conv(plasma_context_t *ctx, const int *dflag, const int *st,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    int flag = 0, word = 0;
    if (plasma_hip_check(hipMemcpyAsync(&flag, dflag, sizeof(int),
                                        hipMemcpyDeviceToHost, ctx->stream),
                         "hipMemcpy(flag)") ||
        plasma_hip_check(hipMemcpyAsync(&word, st, sizeof(int),
                                        hipMemcpyDeviceToHost, ctx->stream),
                         "hipMemcpy(status)") ||
        plasma_hip_check(hipStreamSynchronize(ctx->stream), "sync")) {
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return -1;
    }
    if (word != 0) {
        plasma_request_fail(sequence, request, word);
        return -1;
    }
    return flag ? 1 : 0;
}

/* FP32 solve with the factor in As, in its repeated-solve form when that
 * could be allocated (plasma_rsolve), else by the plain sweeps */
static void solve_s(plasma_enum_t uplo, plasma_desc_t As, plasma_desc_t Xs,
                    const void *inv_s, const void *mpan,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    plasma_rsolve(uplo, As, Xs, inv_s, mpan, sequence, request);
}

/* R = B - A X, the column max-norms of X and R and the convergence flag
 * (reference zcposv.c:466-476 and :509-516), one profiler record (class
 * mixed: the residual's 2 n^2 nrhs flops and its read of A's triangle). */
static int residual_step(plasma_enum_t uplo, plasma_desc_t A,
                         plasma_desc_t B, plasma_desc_t X, plasma_desc_t R,
                         void *symm_work, size_t symm_bytes, double *xnorm,
                         double *rnorm, const double *anorm, double eps_sqrtn,
                         int *dflag, int *st, hipStream_t s,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request)
{
    const double n = A.n, nrhs = X.n;
    const double cf = plasma_is_complex(A.precision) ? 4.0 : 1.0;
    int slot;
    plasma_prof_begin(s, PlasmaProfMixed, cf * 2.0 * n * n * nrhs,
                      n * (n + 1) / 2 * 8.0 + 3.0 * n * nrhs * 8.0, &slot);
    int err =
        issue(plasma_kern_lacpy(B, R, st, s), sequence, request) ||
        issue(plasma_kern_symm_residual(uplo, A, X, R, symm_work, symm_bytes,
                                        st, s), sequence, request) ||
        issue(plasma_kern_amax_cols(X, xnorm, st, s), sequence, request) ||
        issue(plasma_kern_amax_cols(R, rnorm, st, s), sequence, request) ||
        issue(plasma_kern_conv_test(rnorm, xnorm, (int)nrhs, anorm,
                                    eps_sqrtn, dflag, st, s),
              sequence, request);
    plasma_prof_end(s, slot);
    return err;
}

/* FP64 <-> FP32 conversion of a descriptor, one profiler record (class
 * mixed: 12 bytes per element). */
static int lag2_step(plasma_desc_t S, plasma_desc_t D, int *st, hipStream_t s,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    int slot;
    plasma_prof_begin(s, PlasmaProfMixed, 0,
                      (double)plasma_desc_storage_bytes(S) +
                          (double)plasma_desc_storage_bytes(D), &slot);
    int err = issue(plasma_kern_lag2(S, D, st, s), sequence, request);
    plasma_prof_end(s, slot);
    return err;
}

/* hi / lo: the working and the factor precision (RealDouble / RealFloat
 * or ComplexDouble / ComplexFloat) */
static void omp_mixed(plasma_enum_t hi, plasma_enum_t lo, plasma_enum_t uplo,
                      plasma_desc_t A, plasma_desc_t B, plasma_desc_t X,
                      plasma_desc_t As, plasma_desc_t Xs, plasma_desc_t R,
                      double *W, double *Rnorm, double *Xnorm, int *iter,
                      plasma_sequence_t *sequence, plasma_request_t *request)
{
    const int itermax = 30;
    (void)W;
    (void)Rnorm;
    (void)Xnorm;
    if (iter != NULL)
        *iter = 0;
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
    const plasma_desc_t *descs[6] = {&A, &B, &X, &As, &Xs, &R};
    const char *names[6] = {"invalid A", "invalid B", "invalid X",
                            "invalid As", "invalid Xs", "invalid R"};
    for (int i = 0; i < 6; i++) {
        if (plasma_desc_check(*descs[i]) != PlasmaSuccess) {
            plasma_error(names[i]);
            plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
            return;
        }
    }
    if (sequence == NULL || request == NULL || iter == NULL) {
        plasma_error("NULL sequence, request or iter");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (A.n == 0 || B.n == 0)
        return;
    if (A.precision != hi || B.precision != hi || X.precision != hi ||
        R.precision != hi || As.precision != lo || Xs.precision != lo ||
        A.mb != A.nb || As.mb != A.mb || B.mb != A.mb || X.mb != A.mb ||
        R.mb != A.mb || Xs.mb != A.mb || A.m != A.n || B.m != A.m ||
        X.m != A.m || R.m != A.m || Xs.m != A.m || As.m != A.m ||
        B.n != X.n || R.n != X.n || Xs.n != X.n) {
        plasma_error("descriptors do not match the mixed solver's contract");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (sequence->status != PlasmaSuccess)
        return;

    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s = ctx->stream;
    const int n = A.n, nrhs = X.n;
    /* per-call device scratch: norm work | xnorm | rnorm | anorm | flag */
    const size_t lw = plasma_lansy_work_doubles(A);
    double *work = plasma_dev_alloc((lw + 2 * (size_t)nrhs + 2) *
                                    sizeof(double));
    if (work == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    void *inv_s = plasma_dev_alloc(plasma_inv_bytes(As));
    if (inv_s == NULL) {
        plasma_dev_free(work);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    void *inv_d = NULL, *mpan = NULL;
    /* split-K partials of the residual (optional: NULL runs unsplit) */
    const size_t symm_bytes = 8 * (size_t)n * nrhs * sizeof(double);
    void *symm_work = plasma_dev_alloc(symm_bytes);
    double *xnorm = work + lw;
    double *rnorm = xnorm + nrhs;
    double *anorm = rnorm + nrhs;
    int *dflag = (int *)(anorm + 1);
    const double eps = ldexp(1.0, -53);
    const double eps_sqrtn = eps * sqrt((double)n);
    int conv;

    /* reference zcposv.c:442-463 */
    {
        int slot;
        plasma_prof_begin(s, PlasmaProfMixed, 0,
                          (double)plasma_desc_storage_bytes(A), &slot);
        const int err = issue(plasma_kern_lansy_inf(uplo, A, work, anorm, st,
                                                    s), sequence, request);
        plasma_prof_end(s, slot);
        if (err || lag2_step(B, Xs, st, s, sequence, request) ||
            lag2_step(A, As, st, s, sequence, request))
            goto done;
    }
    plasma_pxpotrf(uplo, As, inv_s, sequence, request);
    /* the refinement solves with this factor up to 31 times: prepare its
     * repeated-solve form once (super-block inverses or scaled panels) */
    mpan = plasma_rsolve_prepare(uplo, As, inv_s, sequence, request);
    solve_s(uplo, As, Xs, inv_s, mpan, sequence, request);
    if (lag2_step(Xs, X, st, s, sequence, request))
        goto done;

    /* R = B - A X and the test (reference :466-481) */
    if (residual_step(uplo, A, B, X, R, symm_work, symm_bytes, xnorm, rnorm,
                      anorm, eps_sqrtn, dflag, st, s, sequence, request))
        goto done;
    conv = read_conv(ctx, dflag, st, sequence, request);
    if (conv != 0)
        goto done;  /* converged with iter = 0, or failed */

    /* refinement (reference :485-518) */
    for (int it = 0; it < itermax; it++) {
        if (lag2_step(R, Xs, st, s, sequence, request))
            goto done;
        solve_s(uplo, As, Xs, inv_s, mpan, sequence, request);
        if (lag2_step(Xs, R, st, s, sequence, request) ||
            issue(plasma_kern_geadd(1.0, R, 1.0, X, st, s), sequence,
                  request) ||
            residual_step(uplo, A, B, X, R, symm_work, symm_bytes, xnorm,
                          rnorm, anorm, eps_sqrtn, dflag, st, s, sequence,
                          request))
            goto done;
        conv = read_conv(ctx, dflag, st, sequence, request);
        if (conv < 0)
            goto done;
        if (conv > 0) {
            *iter = it + 1;
            goto done;
        }
    }

    /* no convergence: FP64 factorization (reference :524-538) */
    *iter = -itermax - 1;
    inv_d = plasma_dev_alloc(plasma_inv_bytes(A));
    if (inv_d == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        goto done;
    }
    plasma_pxpotrf(uplo, A, inv_d, sequence, request);
    if (issue(plasma_kern_lacpy(B, X, st, s), sequence, request))
        goto done;
    plasma_pxpotrs(uplo, A, X, inv_d, sequence, request);

done:
    plasma_dev_free(symm_work);
    plasma_dev_free(mpan);
    plasma_dev_free(inv_d);
    plasma_dev_free(inv_s);
    plasma_dev_free(work);
}

void plasma_omp_dsposv(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                       plasma_desc_t X, plasma_desc_t As, plasma_desc_t Xs,
                       plasma_desc_t R, double *W, double *Rnorm,
                       double *Xnorm, int *iter, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    omp_mixed(PlasmaRealDouble, PlasmaRealFloat, uplo, A, B, X, As, Xs, R, W,
              Rnorm, Xnorm, iter, sequence, request);
}

void plasma_omp_zcposv(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                       plasma_desc_t X, plasma_desc_t As, plasma_desc_t Xs,
                       plasma_desc_t R, double *W, double *Rnorm,
                       double *Xnorm, int *iter, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    omp_mixed(PlasmaComplexDouble, PlasmaComplexFloat, uplo, A, B, X, As, Xs,
              R, W, Rnorm, Xnorm, iter, sequence, request);
}

static int mixed_driver(plasma_enum_t hi, plasma_enum_t lo,
                        plasma_enum_t uplo, int n, int nrhs, void *pA,
                        int lda, void *pB, int ldb, void *pX, int ldx,
                        int *iter)
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
    if (ldb < imax(1, n)) {
        plasma_error("illegal value of ldb");
        return -7;
    }
    if (ldx < imax(1, n)) {
        plasma_error("illegal value of ldx");
        return -9;
    }
    *iter = 0;
    if (imin(n, nrhs) == 0)
        return PlasmaSuccess;

    /* reference zcposv.c:160-161: tuned with the single-precision table */
    if (ctx->tuning)
        plasma_tune_potrf(ctx, lo, n);
    const int nb = ctx->nb;
    plasma_desc_t A, B, X, R, As, Xs;
    plasma_desc_t *all[6] = {&A, &B, &X, &R, &As, &Xs};
    int created = 0, retval;
    retval = plasma_desc_triangular_create(hi, uplo, nb, nb, n, n, 0, 0, n,
                                           n, &A);
    if (retval == PlasmaSuccess) created++;
    if (retval == PlasmaSuccess) {
        retval = plasma_desc_general_create(hi, nb, nb, n, nrhs, 0, 0, n,
                                            nrhs, &B);
        if (retval == PlasmaSuccess) created++;
    }
    if (retval == PlasmaSuccess) {
        retval = plasma_desc_general_create(hi, nb, nb, n, nrhs, 0, 0, n,
                                            nrhs, &X);
        if (retval == PlasmaSuccess) created++;
    }
    if (retval == PlasmaSuccess) {
        retval = plasma_desc_general_create(hi, nb, nb, n, nrhs, 0, 0, n,
                                            nrhs, &R);
        if (retval == PlasmaSuccess) created++;
    }
    if (retval == PlasmaSuccess) {
        retval = plasma_desc_triangular_create(lo, uplo, nb, nb, n, n, 0, 0,
                                               n, n, &As);
        if (retval == PlasmaSuccess) created++;
    }
    if (retval == PlasmaSuccess) {
        retval = plasma_desc_general_create(lo, nb, nb, n, nrhs, 0, 0, n,
                                            nrhs, &Xs);
        if (retval == PlasmaSuccess) created++;
    }
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_*_create() failed");
        for (int i = 0; i < created; i++)
            plasma_desc_destroy(all[i]);
        return retval;
    }

    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);

    /* LAPACK arrays: device pointers are used in place, host arrays are
     * translated through the staging path of the layout routines */
    plasma_pxlapack2desc(uplo, pA, lda, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, pB, ldb, B, &sequence, &request);
    omp_mixed(hi, lo, uplo, A, B, X, As, Xs, R, NULL, NULL, NULL, iter,
              &sequence, &request);
    plasma_pxdesc2lapack(PlasmaGeneral, X, pX, ldx, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    for (int i = 0; i < 6; i++)
        plasma_desc_destroy(all[i]);
    return status;
}

int plasma_dsposv(plasma_enum_t uplo, int n, int nrhs, double *pA, int lda,
                  double *pB, int ldb, double *pX, int ldx, int *iter)
{
    return mixed_driver(PlasmaRealDouble, PlasmaRealFloat, uplo, n, nrhs, pA,
                        lda, pB, ldb, pX, ldx, iter);
}

int plasma_zcposv(plasma_enum_t uplo, int n, int nrhs,
                  plasma_complex64_t *pA, int lda, plasma_complex64_t *pB,
                  int ldb, plasma_complex64_t *pX, int ldx, int *iter)
{
    return mixed_driver(PlasmaComplexDouble, PlasmaComplexFloat, uplo, n,
                        nrhs, pA, lda, pB, ldb, pX, ldx, iter);
}

/***************************************************************************
 * Precision conversions (reference compute/zlag2c.c:59-231,
 * compute/clag2z.c:58-226).  hi / lo as in the mixed solver; a complex
 * conversion checks both parts against FLT_MAX (LAPACK zlag2c).
 */
static void omp_lag2(plasma_enum_t hi, plasma_enum_t lo, int to_lo,
                     plasma_desc_t A, plasma_desc_t As,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (plasma_desc_check(A) != PlasmaSuccess ||
        plasma_desc_check(As) != PlasmaSuccess || A.precision != hi ||
        As.precision != lo || A.mt != As.mt || A.nt != As.nt ||
        A.mb != As.mb || A.nb != As.nb) {
        plasma_error(to_lo ? "invalid A or As" : "invalid As or A");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (sequence == NULL || request == NULL) {
        plasma_error("NULL sequence or request");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (A.m == 0 || A.n == 0 || sequence->status != PlasmaSuccess)
        return;
    int *st = plasma_sequence_dstatus(sequence, request);
    if (to_lo)
        issue(plasma_kern_lag2(A, As, st, ctx->stream), sequence, request);
    else
        issue(plasma_kern_lag2(As, A, st, ctx->stream), sequence, request);
}

void plasma_omp_dlag2s(plasma_desc_t A, plasma_desc_t As,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_lag2(PlasmaRealDouble, PlasmaRealFloat, 1, A, As, sequence, request);
}

void plasma_omp_slag2d(plasma_desc_t As, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_lag2(PlasmaRealDouble, PlasmaRealFloat, 0, A, As, sequence, request);
}

void plasma_omp_zlag2c(plasma_desc_t A, plasma_desc_t As,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_lag2(PlasmaComplexDouble, PlasmaComplexFloat, 1, A, As, sequence,
             request);
}

void plasma_omp_clag2z(plasma_desc_t As, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    omp_lag2(PlasmaComplexDouble, PlasmaComplexFloat, 0, A, As, sequence,
             request);
}

static int lag2_driver(plasma_enum_t hi, plasma_enum_t lo, int to_lo, int m,
                       int n, void *pA, int lda, void *pAs, int ldas)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    if (m < 0) {
        plasma_error("illegal value of m");
        return -1;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -2;
    }
    /* argument 4 is the source's leading dimension, 6 the target's */
    int ld_src = to_lo ? lda : ldas, ld_dst = to_lo ? ldas : lda;
    if (ld_src < imax(1, m)) {
        plasma_error("illegal value of the source leading dimension");
        return -4;
    }
    if (ld_dst < imax(1, m)) {
        plasma_error("illegal value of the target leading dimension");
        return -6;
    }
    if (imin(m, n) == 0)
        return PlasmaSuccess;
    const int nb = ctx->nb;
    plasma_desc_t A, As;
    int retval = plasma_desc_general_create(hi, nb, nb, m, n, 0, 0, m, n, &A);
    if (retval != PlasmaSuccess)
        return retval;
    retval = plasma_desc_general_create(lo, nb, nb, m, n, 0, 0, m, n, &As);
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    if (to_lo) {
        plasma_pxlapack2desc(PlasmaGeneral, pA, lda, A, &sequence, &request);
        omp_lag2(hi, lo, 1, A, As, &sequence, &request);
        plasma_pxdesc2lapack(PlasmaGeneral, As, pAs, ldas, &sequence,
                             &request);
    }
    else {
        plasma_pxlapack2desc(PlasmaGeneral, pAs, ldas, As, &sequence,
                             &request);
        omp_lag2(hi, lo, 0, A, As, &sequence, &request);
        plasma_pxdesc2lapack(PlasmaGeneral, A, pA, lda, &sequence, &request);
    }
    int status = plasma_sequence_sync(&sequence);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&As);
    return status;
}

int plasma_dlag2s(int m, int n, double *pA, int lda, float *pAs, int ldas)
{
    return lag2_driver(PlasmaRealDouble, PlasmaRealFloat, 1, m, n, pA, lda,
                       pAs, ldas);
}

int plasma_slag2d(int m, int n, float *pAs, int ldas, double *pA, int lda)
{
    return lag2_driver(PlasmaRealDouble, PlasmaRealFloat, 0, m, n, pA, lda,
                       pAs, ldas);
}

int plasma_zlag2c(int m, int n, plasma_complex64_t *pA, int lda,
                  plasma_complex32_t *pAs, int ldas)
{
    return lag2_driver(PlasmaComplexDouble, PlasmaComplexFloat, 1, m, n, pA,
                       lda, pAs, ldas);
}

int plasma_clag2z(int m, int n, plasma_complex32_t *pAs, int ldas,
                  plasma_complex64_t *pA, int lda)
{
    return lag2_driver(PlasmaComplexDouble, PlasmaComplexFloat, 0, m, n, pA,
                       lda, pAs, ldas);
}
