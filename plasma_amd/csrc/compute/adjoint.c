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
 * the closed form is
 *     Abar = 1/2 L^{-T} Q L^{-1},
 * Q the symmetric matrix whose lower triangle is tril(L^T Lbar) (the
 * symmetric form of I. Murray, "Differentiation of the Cholesky
 * decomposition", 2016, eq. 10: Phi(L^T Lbar) + Phi(L^T Lbar)^T = Q, Phi
 * taking the lower triangle with the diagonal halved).  Applied to the
 * whole matrix it costs two triangular sweeps with n right-hand sides, four
 * times the factorization's flops; here it serves each diagonal tile only.
 *
 * The device runs the blocked reverse sweep (Murray 2016, section 4,
 * blocked level-3 form, restated for a right-looking factorization), about
 * twice the factorization's flops.  G, held in place of Lbar, ends as
 * 2 Abar (lower triangle); for k = mt-1 .. 0, with C = L(k+1:, k) and
 * W = L(k,k)^{-1}:
 *     Gbar(k+1:, k) -= sym(G(k+1:, k+1:)) C        trailing update reversed
 *     P = Gbar(k+1:, k) W                          panel solve reversed
 *     Gbar(k,k) -= P^T C;  G(k+1:, k) = P
 *     G(k,k) = W^T Q_k W,  Q_k symmetric with lower triangle
 *              tril(L(k,k)^T tril(Gbar(k,k)))      the tile's own factor
 * sym() reads the stored lower triangle (diagonal as stored); the last
 * line is the closed form on one tile.  Abar = G / 2, mirrored (exact
 * symmetry).  Upper storage is transposed into lower on the way in
 * (L = U^T, Lbar = Ubar^T).
 *
 * potrs adjoint.  X = A^{-1} B with adjoint Xbar: Bbar = A^{-1} Xbar and
 * Abar -= 1/2 (Bbar X^T + X Bbar^T).
 */
#include <string.h>

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

/* Device scratch of the blocked sweep. */
typedef struct {
    void *panel;      /* (mt - 1) tiles of mb x nb: P */
    void *part;       /* ADJ_MAXSPLIT x nb x nb: Gram partial sums */
    void *symw;       /* split-K partials of the symmetric product */
    size_t symw_bytes;
    void *q, *t1;     /* nb x nb: Q_k, Q_k W */
} adj_work_t;

enum { ADJ_MAXSPLIT = 32, ADJ_SYMSPLIT = 8 };

static void adj_work_free(adj_work_t *w)
{
    plasma_dev_free(w->panel);
    plasma_dev_free(w->part);
    plasma_dev_free(w->symw);
    plasma_dev_free(w->q);
    plasma_dev_free(w->t1);
}

static int adj_work_alloc(plasma_desc_t A, adj_work_t *w)
{
    const size_t es = plasma_element_size(A.precision);
    const size_t tile = (size_t)A.mb * A.nb * es;
    memset(w, 0, sizeof(*w));
    w->panel = plasma_dev_alloc(imax(A.mt - 1, 1) * tile);
    w->part = plasma_dev_alloc(ADJ_MAXSPLIT * tile);
    w->symw_bytes = (size_t)ADJ_SYMSPLIT * A.m * A.nb * es;
    w->symw = plasma_dev_alloc(w->symw_bytes);
    w->q = plasma_dev_alloc(tile);
    w->t1 = plasma_dev_alloc(tile);
    if (!w->panel || !w->part || !w->symw || !w->q || !w->t1) {
        adj_work_free(w);
        return PlasmaErrorOutOfMemory;
    }
    return PlasmaSuccess;
}

/* Algorithmic flops of the diagonal-tile inversions and the blocked sweep
 * (symmetric product 2 m^2 kv, panel product and Gram term 2 m kv^2 each,
 * the tile's own adjoint kv^3 / 3 + 4 kv^3). */
static double adjoint_sweep_flops(plasma_desc_t A)
{
    double flops = 0.0;
    for (int k = 0; k < A.mt; k++) {
        const double kv = plasma_tile_nview(A, k);
        const double rest = A.m - (double)k * A.nb - kv;
        flops += kv * kv * kv * (1.0 / 3.0 + 1.0 / 3.0 + 4.0) +
                 2.0 * rest * rest * kv + 4.0 * rest * kv * kv;
    }
    return flops;
}

/* The blocked reverse sweep on lower storage: Lb (general, lower triangle =
 * Lbar) becomes G = 2 Abar (lower triangle). */
static void adjoint_sweep_lower(plasma_desc_t A, plasma_desc_t Lb,
                                const void *inv, adj_work_t *w, int *st,
                                hipStream_t s, plasma_sequence_t *sequence,
                                plasma_request_t *request)
{
    const size_t es = plasma_element_size(A.precision);
    const int nb = A.nb;
    plasma_desc_t Qd;
    if (issue(plasma_desc_general_init(A.precision, w->q, nb, nb, nb, nb, 0,
                                       0, nb, nb, &Qd),
              sequence, request))
        return;
    for (int k = A.mt - 1; k >= 0 && sequence->status == PlasmaSuccess;
         k--) {
        const int k0 = k * nb;
        const int kv = plasma_tile_nview(A, k);
        const void *W = (const char *)inv + (size_t)k * A.mb * A.mb * es;
        if (k < A.mt - 1) {
            const int m1 = k0 + kv, rest = A.m - m1;
            plasma_desc_t Gt = plasma_desc_view(Lb, m1, m1, rest, rest);
            plasma_desc_t Cv = plasma_desc_view(A, m1, k0, rest, kv);
            plasma_desc_t Rv = plasma_desc_view(Lb, m1, k0, rest, kv);
            if (issue(plasma_kern_symm_residual(PlasmaLower, Gt, Cv, Rv,
                                                w->symw, w->symw_bytes, st,
                                                s),
                      sequence, request))
                break;
            if (issue(plasma_kern_adjoint_panel_step(A, Lb, k, W, w->panel,
                                                     w->part, ADJ_MAXSPLIT,
                                                     st, s),
                      sequence, request))
                break;
        }
        plasma_desc_t Lkk = plasma_desc_view(A, k0, k0, kv, kv);
        plasma_desc_t Gkk = plasma_desc_view(Lb, k0, k0, kv, kv);
        plasma_desc_t Qv = plasma_desc_view(Qd, 0, 0, kv, kv);
        if (issue(plasma_kern_ltl_lower(PlasmaLower, Lkk, Gkk, Qv, st, s),
                  sequence, request))
            break;
        if (issue(plasma_kern_adjoint_diag_sandwich(Lb, k, W, w->q, w->t1,
                                                    st, s),
                  sequence, request))
            break;
    }
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
    /* the symmetric product of each step wants tall panels: tiles of at
     * least 256 once n is large */
    const int nb = (n >= 8192) ? imax(ctx->nb, 256) : ctx->nb;
    const size_t es = plasma_element_size(precision);

    /* L (general descriptor, lower triangle read), G (in place of Lbar),
     * S (staging of upper storage, then the symmetric result) */
    plasma_desc_t A, Lb, S;
    int retval = plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n,
                                            n, &A);
    if (retval != PlasmaSuccess)
        return retval;
    retval = plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n,
                                        &Lb);
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        return retval;
    }
    retval = plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n,
                                        &S);
    if (retval != PlasmaSuccess) {
        plasma_desc_destroy(&A);
        plasma_desc_destroy(&Lb);
        return retval;
    }
    void *inv = plasma_dev_alloc(plasma_inv_bytes(A));
    adj_work_t work;
    lapack_arg_t a, ab;
    retval = inv ? adj_work_alloc(A, &work) : PlasmaErrorOutOfMemory;
    if (retval == PlasmaSuccess) {
        retval = plasma_arg_stage_in(&a, (void *)pA, lda, n, n, es, 1);
        if (retval != PlasmaSuccess)
            adj_work_free(&work);
    }
    if (retval == PlasmaSuccess) {
        retval = plasma_arg_stage_in(&ab, pAbar, ldabar, n, n, es, 1);
        if (retval != PlasmaSuccess) {
            plasma_arg_stage_out(&a, 0);
            adj_work_free(&work);
        }
    }
    if (retval != PlasmaSuccess) {
        plasma_dev_free(inv);
        plasma_desc_destroy(&A);
        plasma_desc_destroy(&Lb);
        plasma_desc_destroy(&S);
        return retval;
    }

    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    hipStream_t s = ctx->stream;
    int *st = plasma_sequence_dstatus(&sequence, &request);
    const double dn = n;
    int slot;
    if (uplo == PlasmaLower) {
        plasma_pxlapack2desc(PlasmaGeneral, a.dev, a.ld, A, &sequence,
                             &request);
        plasma_pxlapack2desc(PlasmaGeneral, ab.dev, ab.ld, Lb, &sequence,
                             &request);
    }
    else if (sequence.status == PlasmaSuccess && st != NULL) {
        /* L = U^T and Lbar = Ubar^T: the sweep runs on lower storage */
        plasma_pxlapack2desc(PlasmaGeneral, a.dev, a.ld, S, &sequence,
                             &request);
        plasma_prof_begin(s, PlasmaProfAdjoint, 0, 2.0 * dn * dn * es,
                          &slot);
        issue(plasma_kern_tile_axpby_t(0.0, S, 1.0, S, A, 0, st, s),
              &sequence, &request);
        plasma_prof_end(s, slot);
        plasma_pxlapack2desc(PlasmaGeneral, ab.dev, ab.ld, S, &sequence,
                             &request);
        plasma_prof_begin(s, PlasmaProfAdjoint, 0, 2.0 * dn * dn * es,
                          &slot);
        issue(plasma_kern_tile_axpby_t(0.0, S, 1.0, S, Lb, 0, st, s),
              &sequence, &request);
        plasma_prof_end(s, slot);
    }
    if (sequence.status == PlasmaSuccess && st != NULL) {
        plasma_prof_begin(s, PlasmaProfAdjoint, adjoint_sweep_flops(A), 0,
                          &slot);
        if (!issue(plasma_kern_trtri_diag(PlasmaLower, A, inv, st, s),
                   &sequence, &request))
            adjoint_sweep_lower(A, Lb, inv, &work, st, s, &sequence,
                                &request);
        plasma_prof_end(s, slot);
    }
    if (sequence.status == PlasmaSuccess && st != NULL) {
        /* Abar = G / 2, both triangles from G's lower one */
        plasma_prof_begin(s, PlasmaProfAdjoint, 0, 2.0 * dn * dn * es,
                          &slot);
        issue(plasma_kern_tile_axpby_t(0.5, Lb, 0.5, Lb, S, 1, st, s),
              &sequence, &request);
        plasma_prof_end(s, slot);
    }
    plasma_pxdesc2lapack(PlasmaGeneral, S, ab.dev, ab.ld, &sequence,
                         &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_arg_stage_out(&a, 0);
    plasma_arg_stage_out(&ab, status == PlasmaSuccess);
    plasma_dev_free(inv);
    adj_work_free(&work);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&Lb);
    plasma_desc_destroy(&S);
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
