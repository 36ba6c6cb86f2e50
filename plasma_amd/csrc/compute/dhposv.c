/*
 * dhposv.c -- plasma_dhposv: the FP64 solution of A X = B (A symmetric
 * positive definite) from a bf16-input factor refined by GMRES-IR.  No
 * reference counterpart: the reference's mixed solver is zcposv (FP32
 * factor, classical refinement), restated in dsposv.c.
 *
 * 1. Symmetric diagonal equilibration: S = diag(a_ii^{-1/2}).  The system
 *    solved is (S A S) Y = S B with X = S Y, so the scaled matrix has a
 *    unit diagonal and its entries lie in [-1, 1].
 * 2. The scaled matrix, rounded to FP32, is factored with its trailing
 *    updates on bf16 MFMA from bf16 inputs (one product per FP32 product,
 *    FP32 accumulation; the diagonal tiles and panels in FP32).
 * 3. GMRES-IR (E. Carson and N. J. Higham, "Accelerating the solution of
 *    linear systems by iterative refinement in three precisions", SIAM J.
 *    Sci. Comput. 40 (2018)): every refinement step solves the correction
 *    equation A d = r by FP64 GMRES on M^{-1} A, M = L L^T the factor (left
 *    preconditioning; M^{-1} applied by the FP32 sweeps), restarted after
 *    DH_RESTART inner iterations.  The right-hand sides run in lockstep,
 *    each with its own Krylov basis and Hessenberg matrix: classical
 *    Gram-Schmidt applied twice, Givens rotations on the host.
 * The stopping test is dsposv's (reference zcposv.c:466-481) on the scaled
 * system.  When the factor breaks down (a pivot that is not positive in
 * low precision) or DH_OUTER refinement steps do not converge, the scaled
 * system is factored and solved in FP64.
 *
 * iter: the total number of inner GMRES iterations; -1 after the FP64
 * fallback.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

enum { DH_RESTART = 20, DH_OUTER = 10 };
#define DH_RTOL 1e-10   /* inner GMRES: relative preconditioned residual */

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err)
        plasma_request_fail(sequence, request, err);
    return err;
}

/* Synchronous copy on the main stream (the stream's earlier work first). */
static int copy_sync(plasma_context_t *ctx, void *dst, const void *src,
                     size_t bytes, hipMemcpyKind kind,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (plasma_hip_check(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream),
                         "hipMemcpy") ||
        plasma_hip_check(hipStreamSynchronize(ctx->stream), "sync")) {
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return -1;
    }
    return 0;
}

typedef struct {
    plasma_context_t *ctx;
    plasma_enum_t uplo;
    plasma_desc_t A, As, Xs, T;    /* T: FP64 n x nrhs scratch */
    void *inv_s, *mpan;
    void *symm_work;
    size_t symm_bytes;
    int *st;
    hipStream_t s;
    char *vbase;                   /* DH_RESTART + 1 basis vectors */
    size_t vbytes;
    double *dots, *coef;           /* device: (DH_RESTART + 1) x nrhs */
} dh_t;

static plasma_desc_t vec(const dh_t *d, int i)
{
    plasma_desc_t v = d->T;
    v.matrix = d->vbase + (size_t)i * d->vbytes;
    return v;
}

static plasma_vecset_t vset(const dh_t *d, int first, int count)
{
    plasma_vecset_t v;
    memset(&v, 0, sizeof(v));
    v.nv = count;
    for (int i = 0; i < count; i++)
        v.p[i] = (const double *)vec(d, first + i).matrix;
    return v;
}

static plasma_vecset_t one(const plasma_desc_t *D)
{
    plasma_vecset_t v;
    memset(&v, 0, sizeof(v));
    v.nv = 1;
    v.p[0] = (const double *)D->matrix;
    return v;
}

/* Z = M^{-1} Y in FP32 (Y, Z: FP64 descriptors, may alias) */
static void precond(const dh_t *d, plasma_desc_t Y, plasma_desc_t Z,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (issue(plasma_kern_lag2(Y, d->Xs, d->st, d->s), sequence, request))
        return;
    plasma_rsolve(d->uplo, d->As, d->Xs, d->inv_s, d->mpan, sequence,
                  request);
    if (sequence->status == PlasmaSuccess)
        issue(plasma_kern_lag2(d->Xs, Z, d->st, d->s), sequence, request);
}

/* One restarted GMRES cycle on A d = R from d = 0, added into X; returns
 * the inner iterations taken (0 on failure). */
static int gmres_cycle(const dh_t *d, plasma_desc_t R, plasma_desc_t X,
                       plasma_sequence_t *sequence, plasma_request_t *request)
{
    const int nrhs = X.n, m = DH_RESTART;
    const int *st = d->st;
    hipStream_t s = d->s;
    double *host = malloc(sizeof(double) *
                          ((size_t)(m + 1) * nrhs * 2 +      /* h, y */
                           (size_t)nrhs * (m + 1) * m +      /* Rmat */
                           (size_t)nrhs * m * 2 +            /* cs, sn */
                           (size_t)nrhs * (m + 1) +          /* gamma */
                           (size_t)nrhs * 2));               /* beta, inv */
    if (host == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return 0;
    }
    double *h = host, *y = h + (size_t)(m + 1) * nrhs;
    double *Rm = y + (size_t)(m + 1) * nrhs;
    double *cs = Rm + (size_t)nrhs * (m + 1) * m, *sn = cs + (size_t)nrhs * m;
    double *gam = sn + (size_t)nrhs * m;
    double *beta = gam + (size_t)nrhs * (m + 1), *inv = beta + nrhs;
    int jj = 0;
    plasma_desc_t Z = d->T;

    /* v_0 = -M^{-1} R / beta: the operator below is -M^{-1} A */
    precond(d, R, Z, sequence, request);
    plasma_vecset_t vz = one(&Z);
    if (sequence->status != PlasmaSuccess ||
        issue(plasma_kern_coldots(Z, vz, d->dots, st, s), sequence,
              request) ||
        copy_sync(d->ctx, beta, d->dots, sizeof(double) * nrhs,
                  hipMemcpyDeviceToHost, sequence, request))
        goto out;
    for (int c = 0; c < nrhs; c++) {
        beta[c] = sqrt(beta[c]);
        inv[c] = beta[c] > 0.0 ? 1.0 / beta[c] : 0.0;
        gam[(size_t)c * (m + 1)] = beta[c];
    }
    if (copy_sync(d->ctx, d->coef, inv, sizeof(double) * nrhs,
                  hipMemcpyHostToDevice, sequence, request) ||
        issue(plasma_kern_colcombine(vec(d, 0), 0.0, -1.0, vz, d->coef, st,
                                     s),
              sequence, request))
        goto out;

    for (int j = 0; j < m; j++) {
        /* Z = -M^{-1} A v_j */
        if (issue(plasma_kern_colcombine(Z, 0.0, 0.0, vset(d, 0, 0), NULL,
                                         st, s),
                  sequence, request) ||
            issue(plasma_kern_symm_residual(d->uplo, d->A, vec(d, j), Z,
                                            d->symm_work, d->symm_bytes, st,
                                            s),
                  sequence, request))
            goto out;
        precond(d, Z, Z, sequence, request);
        if (sequence->status != PlasmaSuccess)
            goto out;
        /* classical Gram-Schmidt, twice: h = V^T Z, Z -= V h */
        const plasma_vecset_t V = vset(d, 0, j + 1);
        for (int c = 0; c <= j; c++)
            for (int r = 0; r < nrhs; r++)
                h[(size_t)c * nrhs + r] = 0.0;
        for (int pass = 0; pass < 2; pass++) {
            if (issue(plasma_kern_coldots(Z, V, d->dots, st, s), sequence,
                      request) ||
                issue(plasma_kern_colcombine(Z, 1.0, -1.0, V, d->dots, st,
                                             s),
                      sequence, request) ||
                copy_sync(d->ctx, y, d->dots,
                          sizeof(double) * (size_t)(j + 1) * nrhs,
                          hipMemcpyDeviceToHost, sequence, request))
                goto out;
            for (size_t e = 0; e < (size_t)(j + 1) * nrhs; e++)
                h[e] += y[e];
        }
        /* h_{j+1,j} = ||Z||, v_{j+1} = Z / h_{j+1,j} */
        if (issue(plasma_kern_coldots(Z, vz, d->dots, st, s), sequence,
                  request) ||
            copy_sync(d->ctx, y, d->dots, sizeof(double) * nrhs,
                      hipMemcpyDeviceToHost, sequence, request))
            goto out;
        int done = 1;
        for (int c = 0; c < nrhs; c++) {
            const double hn = sqrt(y[c]);
            inv[c] = hn > 0.0 ? 1.0 / hn : 0.0;
            /* column j of the Hessenberg matrix of right-hand side c */
            double hc[DH_RESTART + 1];
            for (int i = 0; i <= j; i++)
                hc[i] = h[(size_t)i * nrhs + c];
            hc[j + 1] = hn;
            double *csc = cs + (size_t)c * m, *snc = sn + (size_t)c * m;
            for (int i = 0; i < j; i++) {
                const double t = csc[i] * hc[i] + snc[i] * hc[i + 1];
                hc[i + 1] = -snc[i] * hc[i] + csc[i] * hc[i + 1];
                hc[i] = t;
            }
            const double rr = hypot(hc[j], hc[j + 1]);
            csc[j] = rr > 0.0 ? hc[j] / rr : 1.0;
            snc[j] = rr > 0.0 ? hc[j + 1] / rr : 0.0;
            hc[j] = rr;
            double *g = gam + (size_t)c * (m + 1);
            g[j + 1] = -snc[j] * g[j];
            g[j] = csc[j] * g[j];
            double *Rc = Rm + (size_t)c * (m + 1) * m;
            for (int i = 0; i <= j; i++)
                Rc[(size_t)j * (m + 1) + i] = hc[i];
            if (fabs(g[j + 1]) > DH_RTOL * beta[c])
                done = 0;
        }
        jj = j + 1;
        if (done || j + 1 == m)
            break;
        if (copy_sync(d->ctx, d->coef, inv, sizeof(double) * nrhs,
                      hipMemcpyHostToDevice, sequence, request) ||
            issue(plasma_kern_colcombine(vec(d, j + 1), 0.0, 1.0, vz,
                                         d->coef, st, s),
                  sequence, request))
            goto out;
    }

    /* y = R^{-1} gamma per right-hand side, X += V y */
    for (int c = 0; c < nrhs; c++) {
        const double *Rc = Rm + (size_t)c * (m + 1) * m;
        const double *g = gam + (size_t)c * (m + 1);
        for (int i = jj - 1; i >= 0; i--) {
            double t = g[i];
            for (int l = i + 1; l < jj; l++)
                t -= Rc[(size_t)l * (m + 1) + i] * y[(size_t)l * nrhs + c];
            const double rii = Rc[(size_t)i * (m + 1) + i];
            y[(size_t)i * nrhs + c] = rii != 0.0 ? t / rii : 0.0;
        }
    }
    if (copy_sync(d->ctx, d->coef, y, sizeof(double) * (size_t)jj * nrhs,
                  hipMemcpyHostToDevice, sequence, request) ||
        issue(plasma_kern_colcombine(X, 1.0, 1.0, vset(d, 0, jj), d->coef,
                                     st, s),
              sequence, request))
        jj = 0;
out:
    free(host);
    return sequence->status == PlasmaSuccess ? jj : 0;
}

static int read_word(plasma_context_t *ctx, const int *dword, int *word,
                     plasma_sequence_t *sequence, plasma_request_t *request)
{
    *word = 0;
    return copy_sync(ctx, word, dword, sizeof(int), hipMemcpyDeviceToHost,
                     sequence, request);
}

/* The stopping test of dsposv on the scaled system: R = B - A X, then
 * flag = every column satisfies ||r||_inf <= ||x||_inf ||A||_inf eps
 * sqrt(n).  Returns 1 (converged), 0, or -1 (failure). */
static int residual_test(const dh_t *d, plasma_desc_t B, plasma_desc_t X,
                         plasma_desc_t R, double *xnorm, double *rnorm,
                         const double *anorm, int *dflag, double eps_sqrtn,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request)
{
    const int nrhs = X.n;
    if (issue(plasma_kern_lacpy(B, R, d->st, d->s), sequence, request) ||
        issue(plasma_kern_symm_residual(d->uplo, d->A, X, R, d->symm_work,
                                        d->symm_bytes, d->st, d->s),
              sequence, request) ||
        issue(plasma_kern_amax_cols(X, xnorm, d->st, d->s), sequence,
              request) ||
        issue(plasma_kern_amax_cols(R, rnorm, d->st, d->s), sequence,
              request) ||
        issue(plasma_kern_conv_test(rnorm, xnorm, nrhs, anorm, eps_sqrtn,
                                    dflag, d->st, d->s),
              sequence, request))
        return -1;
    int flag = 0;
    if (copy_sync(d->ctx, &flag, dflag, sizeof(int), hipMemcpyDeviceToHost,
                  sequence, request))
        return -1;
    return flag ? 1 : 0;
}

int plasma_dhposv(plasma_enum_t uplo, int n, int nrhs, double *pA, int lda,
                  double *pB, int ldb, double *pX, int ldx, int *iter)
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
    if (iter == NULL) {
        plasma_error("NULL iter");
        return -10;
    }
    *iter = 0;
    if (imin(n, nrhs) == 0)
        return PlasmaSuccess;

    if (ctx->tuning)
        plasma_tune_potrf(ctx, PlasmaRealFloat, n);
    const int nb = ctx->nb;
    plasma_desc_t A, B, X, R, T, As, Xs;
    plasma_desc_t *all[7] = {&A, &B, &X, &R, &T, &As, &Xs};
    int created = 0, retval;
    retval = plasma_desc_triangular_create(PlasmaRealDouble, uplo, nb, nb, n,
                                           n, 0, 0, n, n, &A);
    if (retval == PlasmaSuccess) created++;
    for (int i = 1; i < 5 && retval == PlasmaSuccess; i++) {
        retval = plasma_desc_general_create(PlasmaRealDouble, nb, nb, n,
                                            nrhs, 0, 0, n, nrhs, all[i]);
        if (retval == PlasmaSuccess) created++;
    }
    if (retval == PlasmaSuccess) {
        retval = plasma_desc_triangular_create(PlasmaRealFloat, uplo, nb, nb,
                                               n, n, 0, 0, n, n, &As);
        if (retval == PlasmaSuccess) created++;
    }
    if (retval == PlasmaSuccess) {
        retval = plasma_desc_general_create(PlasmaRealFloat, nb, nb, n, nrhs,
                                            0, 0, n, nrhs, &Xs);
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
    plasma_pxlapack2desc(uplo, pA, lda, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, pB, ldb, B, &sequence, &request);

    dh_t d;
    memset(&d, 0, sizeof(d));
    d.ctx = ctx;
    d.uplo = uplo;
    d.A = A;
    d.As = As;
    d.Xs = Xs;
    d.T = T;
    d.s = ctx->stream;
    d.st = plasma_sequence_dstatus(&sequence, &request);
    d.symm_bytes = 8 * (size_t)n * nrhs * sizeof(double);
    d.vbytes = plasma_desc_storage_bytes(T);
    const size_t lw = plasma_lansy_work_doubles(A);
    double *work = plasma_dev_alloc((lw + 2 * (size_t)nrhs + 2 + n) *
                                    sizeof(double));
    d.inv_s = plasma_dev_alloc(plasma_inv_bytes(As));
    d.symm_work = plasma_dev_alloc(d.symm_bytes);
    d.vbase = plasma_dev_alloc(d.vbytes * (DH_RESTART + 1));
    d.dots = plasma_dev_alloc(2 * sizeof(double) * (DH_RESTART + 1) * nrhs);
    void *inv_d = NULL;
    int fallback = 0, total = 0, word = 0;
    if (work == NULL || d.inv_s == NULL || d.vbase == NULL ||
        d.dots == NULL) {
        plasma_request_fail(&sequence, &request, PlasmaErrorOutOfMemory);
        goto done;
    }
    d.coef = d.dots + (size_t)(DH_RESTART + 1) * nrhs;
    double *xnorm = work + lw, *rnorm = xnorm + nrhs, *anorm = rnorm + nrhs;
    int *dflag = (int *)(anorm + 1);
    double *scale = anorm + 2;
    const double eps_sqrtn = ldexp(1.0, -53) * sqrt((double)n);
    if (sequence.status != PlasmaSuccess || d.st == NULL)
        goto done;

    /* S A S, S B, ||S A S||_inf, the FP32 copy */
    if (issue(plasma_kern_equil_diag(uplo, A, scale, d.st, d.s), &sequence,
              &request) ||
        issue(plasma_kern_scale_sym(uplo, A, scale, d.st, d.s), &sequence,
              &request) ||
        issue(plasma_kern_scale_rows(B, scale, d.st, d.s), &sequence,
              &request) ||
        issue(plasma_kern_lansy_inf(uplo, A, work, anorm, d.st, d.s),
              &sequence, &request) ||
        issue(plasma_kern_lag2(A, As, d.st, d.s), &sequence, &request))
        goto done;

    /* the bf16-input factor; a breakdown clears the status word and falls
     * back to FP64 */
    ctx->low_factor = ctx->half_input == PlasmaSplitBf16 ? PlasmaLowBf16
                                                         : PlasmaLowFp16;
    plasma_pxpotrf(uplo, As, d.inv_s, &sequence, &request);
    ctx->low_factor = 0;
    if (sequence.status != PlasmaSuccess ||
        read_word(ctx, d.st, &word, &sequence, &request))
        goto done;
    if (word != 0) {
        const int zero = 0;
        if (copy_sync(ctx, d.st, &zero, sizeof(int), hipMemcpyHostToDevice,
                      &sequence, &request))
            goto done;
        fallback = 1;
    }
    else {
        d.mpan = plasma_rsolve_prepare(uplo, As, d.inv_s, &sequence,
                                       &request);
        /* X = M^{-1} S B, then refinement steps */
        precond(&d, B, X, &sequence, &request);
        int conv = 0;
        for (int outer = 0; outer <= DH_OUTER && sequence.status ==
                                                     PlasmaSuccess;
             outer++) {
            conv = residual_test(&d, B, X, R, xnorm, rnorm, anorm, dflag,
                                 eps_sqrtn, &sequence, &request);
            if (conv != 0 || outer == DH_OUTER)
                break;
            const int k = gmres_cycle(&d, R, X, &sequence, &request);
            if (k == 0)
                break;
            total += k;
        }
        if (sequence.status != PlasmaSuccess)
            goto done;
        fallback = conv != 1;
    }

    if (fallback) {
        /* the scaled system in FP64 */
        inv_d = plasma_dev_alloc(plasma_inv_bytes(A));
        if (inv_d == NULL) {
            plasma_request_fail(&sequence, &request, PlasmaErrorOutOfMemory);
            goto done;
        }
        plasma_pxpotrf(uplo, A, inv_d, &sequence, &request);
        if (issue(plasma_kern_lacpy(B, X, d.st, d.s), &sequence, &request))
            goto done;
        plasma_pxpotrs(uplo, A, X, inv_d, &sequence, &request);
    }
    /* X = S Y */
    issue(plasma_kern_scale_rows(X, scale, d.st, d.s), &sequence, &request);
    *iter = fallback ? -1 : total;

done:
    plasma_pxdesc2lapack(PlasmaGeneral, X, pX, ldx, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    ctx->low_factor = 0;
    plasma_dev_free(inv_d);
    plasma_dev_free(d.mpan);
    plasma_dev_free(d.dots);
    plasma_dev_free(d.vbase);
    plasma_dev_free(d.symm_work);
    plasma_dev_free(d.inv_s);
    plasma_dev_free(work);
    for (int i = 0; i < 7; i++)
        plasma_desc_destroy(all[i]);
    return status;
}
