/*
 * blas3.c -- the public tile BLAS-3 front ends of the Cholesky path:
 * plasma_{d,s}gemm, plasma_{d,s}syrk, plasma_{d,s}trsm and their
 * plasma_omp_* tile forms.
 *
 * Signatures, argument checks, return codes and quick returns follow the
 * reference drivers (compute/zgemm.c:105-244 and :315-380, compute/
 * zsyrk.c:89-206 and :263-323, compute/ztrsm.c:104-225 and :290-360): the
 * LAPACK arrays are translated into general tile descriptors of the
 * context's nb, the tile routine runs, and the output is translated back.
 * For real precisions PlasmaConjTrans means PlasmaTrans.
 *
 *  * gemm / syrk: one MFMA launch over the 128 x 128 blocks of C, each
 *    block summing all K tiles in its accumulators (the reference's
 *    pzgemm / pzsyrk issue one tile task per (m, n, k)).
 *  * trsm (reference compute/pztrsm.c, all side / uplo / trans / diag
 *    variants): the diagonal tiles of A are inverted once (W_k, lower-
 *    equivalent), then a right-looking block sweep: per step k the solved
 *    block row (or column) T = op(W_k) B_k (one MFMA launch), the update of
 *    the remaining block rows (columns) by T (one launch), and the copy of
 *    T into B.  With L = A (lower) or A^T (upper) and W_k = L_kk^{-1},
 *    op(A_kk)^{-1} = W_k for (Lower, NoTrans) and (Upper, Trans) and W_k^T
 *    for the other two.  alpha scales B first.
 */
#include "internal.h"

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err)
        plasma_request_fail(sequence, request, err);
    return err;
}

static int is_trans(plasma_enum_t t)
{
    return t == PlasmaTrans || t == PlasmaConjTrans;
}

static int common_checks(plasma_sequence_t *sequence,
                         plasma_request_t *request)
{
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
    return 1;
}

static int desc_ok(plasma_desc_t X, const char *what,
                   plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (plasma_desc_check(X) != PlasmaSuccess) {
        plasma_error(what);
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    return 1;
}

/* The kernels need one square tile size and precision throughout. */
static int same_tiling(plasma_enum_t precision, plasma_desc_t A,
                       plasma_desc_t B, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    if (A.precision != precision || B.precision != precision ||
        A.mb != A.nb || B.mb != A.mb || B.nb != A.nb) {
        plasma_error("descriptors must share the routine's precision and "
                     "one square tile size");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return 0;
    }
    return 1;
}

/***************************************************************************
 * gemm
 */
void plasma_omp_xgemm(plasma_enum_t precision, plasma_enum_t transa,
                      plasma_enum_t transb, double alpha, plasma_desc_t A,
                      plasma_desc_t B, double beta, plasma_desc_t C,
                      plasma_sequence_t *sequence, plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (transa != PlasmaNoTrans && !is_trans(transa)) {
        plasma_error("illegal value of transa");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (transb != PlasmaNoTrans && !is_trans(transb)) {
        plasma_error("illegal value of transb");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!desc_ok(A, "invalid A", sequence, request) ||
        !desc_ok(B, "invalid B", sequence, request) ||
        !desc_ok(C, "invalid C", sequence, request) ||
        !common_checks(sequence, request))
        return;
    const int k = transa == PlasmaNoTrans ? A.n : A.m;
    if (C.m == 0 || C.n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0))
        return;
    if (!same_tiling(precision, A, C, sequence, request) ||
        !same_tiling(precision, B, C, sequence, request))
        return;
    const int am = transa == PlasmaNoTrans ? A.m : A.n;
    const int bk = transb == PlasmaNoTrans ? B.m : B.n;
    const int bn = transb == PlasmaNoTrans ? B.n : B.m;
    if (am != C.m || bn != C.n || bk != k) {
        plasma_error("operand shapes do not match C");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (sequence->status != PlasmaSuccess)
        return;
    int *st = plasma_sequence_dstatus(sequence, request);
    int slot;
    plasma_prof_begin(ctx->stream, PlasmaProfUpdate,
                      ctx->profile ? 2.0 * C.m * C.n * (double)k : 0, 0,
                      &slot);
    issue(plasma_kern_gemm(transa, transb, alpha, A, B, beta, C,
                           PlasmaGeneral, st, ctx->stream),
          sequence, request);
    plasma_prof_end(ctx->stream, slot);
}

int plasma_xgemm(plasma_enum_t precision, plasma_enum_t transa,
                 plasma_enum_t transb, int m, int n, int k, double alpha,
                 void *pA, int lda, void *pB, int ldb, double beta, void *pC,
                 int ldc)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    if (transa != PlasmaNoTrans && !is_trans(transa)) {
        plasma_error("illegal value of transa");
        return -1;
    }
    if (transb != PlasmaNoTrans && !is_trans(transb)) {
        plasma_error("illegal value of transb");
        return -2;
    }
    if (m < 0) {
        plasma_error("illegal value of m");
        return -3;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -4;
    }
    if (k < 0) {
        plasma_error("illegal value of k");
        return -5;
    }
    const int am = transa == PlasmaNoTrans ? m : k;
    const int an = transa == PlasmaNoTrans ? k : m;
    const int bm = transb == PlasmaNoTrans ? k : n;
    const int bn = transb == PlasmaNoTrans ? n : k;
    if (lda < imax(1, am)) {
        plasma_error("illegal value of lda");
        return -8;
    }
    if (ldb < imax(1, bm)) {
        plasma_error("illegal value of ldb");
        return -10;
    }
    if (ldc < imax(1, m)) {
        plasma_error("illegal value of ldc");
        return -13;
    }
    if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0))
        return PlasmaSuccess;

    const int nb = ctx->nb;
    plasma_desc_t A, B, C;
    int retval = plasma_desc_general_create(precision, nb, nb, am, an, 0, 0,
                                            am, an, &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        return retval;
    }
    retval = plasma_desc_general_create(precision, nb, nb, bm, bn, 0, 0, bm,
                                        bn, &B);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        plasma_desc_destroy(&A);
        return retval;
    }
    retval = plasma_desc_general_create(precision, nb, nb, m, n, 0, 0, m, n,
                                        &C);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        plasma_desc_destroy(&A);
        plasma_desc_destroy(&B);
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    plasma_pxlapack2desc(PlasmaGeneral, pA, lda, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, pB, ldb, B, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, pC, ldc, C, &sequence, &request);
    plasma_omp_xgemm(precision, transa, transb, alpha, A, B, beta, C,
                     &sequence, &request);
    plasma_pxdesc2lapack(PlasmaGeneral, C, pC, ldc, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&B);
    plasma_desc_destroy(&C);
    return status;
}

/***************************************************************************
 * syrk
 */
void plasma_omp_xsyrk(plasma_enum_t precision, plasma_enum_t uplo,
                      plasma_enum_t trans, double alpha, plasma_desc_t A,
                      double beta, plasma_desc_t C,
                      plasma_sequence_t *sequence, plasma_request_t *request)
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
    if (trans != PlasmaNoTrans && trans != PlasmaTrans) {
        plasma_error("illegal value of trans");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!desc_ok(A, "invalid A", sequence, request) ||
        !desc_ok(C, "invalid C", sequence, request) ||
        !common_checks(sequence, request))
        return;
    const int k = trans == PlasmaNoTrans ? A.n : A.m;
    if (C.m == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0))
        return;
    if (!same_tiling(precision, A, C, sequence, request))
        return;
    const int an = trans == PlasmaNoTrans ? A.m : A.n;
    if (C.m != C.n || an != C.m) {
        plasma_error("A does not match the order of C");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (sequence->status != PlasmaSuccess)
        return;
    int *st = plasma_sequence_dstatus(sequence, request);
    /* C = alpha A A^T (NoTrans) or alpha A^T A (Trans), one triangle */
    const plasma_enum_t ta = trans, tb = trans == PlasmaNoTrans
                                             ? PlasmaTrans : PlasmaNoTrans;
    int slot;
    plasma_prof_begin(ctx->stream, PlasmaProfUpdate,
                      ctx->profile ? (double)C.m * (C.m + 1) * k : 0, 0,
                      &slot);
    issue(plasma_kern_gemm(ta, tb, alpha, A, A, beta, C, uplo, st,
                           ctx->stream),
          sequence, request);
    plasma_prof_end(ctx->stream, slot);
}

int plasma_xsyrk(plasma_enum_t precision, plasma_enum_t uplo,
                 plasma_enum_t trans, int n, int k, double alpha, void *pA,
                 int lda, double beta, void *pC, int ldc)
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
    if (trans != PlasmaNoTrans && trans != PlasmaTrans) {
        plasma_error("illegal value of trans");
        return -2;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -3;
    }
    if (k < 0) {
        plasma_error("illegal value of k");
        return -4;
    }
    const int am = trans == PlasmaNoTrans ? n : k;
    const int an = trans == PlasmaNoTrans ? k : n;
    if (lda < imax(1, am)) {
        plasma_error("illegal value of lda");
        return -7;
    }
    if (ldc < imax(1, n)) {
        plasma_error("illegal value of ldc");
        return -10;
    }
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0))
        return PlasmaSuccess;

    const int nb = ctx->nb;
    plasma_desc_t A, C;
    int retval = plasma_desc_general_create(precision, nb, nb, am, an, 0, 0,
                                            am, an, &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        return retval;
    }
    retval = plasma_desc_general_create(precision, nb, nb, n, n, 0, 0, n, n,
                                        &C);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        plasma_desc_destroy(&A);
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    plasma_pxlapack2desc(PlasmaGeneral, pA, lda, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, pC, ldc, C, &sequence, &request);
    plasma_omp_xsyrk(precision, uplo, trans, alpha, A, beta, C, &sequence,
                     &request);
    plasma_pxdesc2lapack(PlasmaGeneral, C, pC, ldc, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&C);
    return status;
}

/***************************************************************************
 * trsm
 */
/* Tile-aligned view of rows [r0, r1) x columns [c0, c1) (tile indices). */
static plasma_desc_t tview(plasma_desc_t X, int r0, int r1, int c0, int c1)
{
    const int i = r0 * X.mb, j = c0 * X.nb;
    const int m = imin(r1 * X.mb, X.m) - i, n = imin(c1 * X.nb, X.n) - j;
    return plasma_desc_view(X, i, j, m, n);
}

static void pxtrsm(plasma_enum_t precision, plasma_enum_t side,
                   plasma_enum_t uplo, plasma_enum_t transa,
                   plasma_enum_t diag, double alpha, plasma_desc_t A,
                   plasma_desc_t B, plasma_sequence_t *sequence,
                   plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    hipStream_t s = ctx->stream;
    int *st = plasma_sequence_dstatus(sequence, request);
    const int nb = A.nb, at = A.mt;
    const int left = side == PlasmaLeft, ta = is_trans(transa);
    /* op(A) lower triangular: sweep forward for Left, backward for Right */
    const int op_lower = (uplo == PlasmaLower) != ta;
    const int forward = left ? op_lower : !op_lower;
    const plasma_enum_t wop = ((uplo == PlasmaUpper) != ta) ? PlasmaTrans
                                                            : PlasmaNoTrans;

    /* D: the diagonal tiles as lower triangles; W: their inverses (tile
     * (0, k) of a one-tile-row descriptor each) */
    plasma_desc_t D, W, T;
    if (plasma_desc_general_create(precision, nb, nb, nb, A.n, 0, 0, nb, A.n,
                                   &D) != PlasmaSuccess) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    if (plasma_desc_general_create(precision, nb, nb, nb, at * nb, 0, 0, nb,
                                   at * nb, &W) != PlasmaSuccess) {
        plasma_desc_destroy(&D);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    /* T: one solved block row (Left) or block column (Right) of B */
    const int tm = left ? nb : B.m, tn = left ? B.n : nb;
    if (plasma_desc_general_create(precision, nb, nb, tm, tn, 0, 0, tm, tn,
                                   &T) != PlasmaSuccess) {
        plasma_desc_destroy(&D);
        plasma_desc_destroy(&W);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }

    int err = plasma_kern_trsm_diag_prep(uplo, diag, A, D, st, s);
    if (!err)
        err = plasma_kern_trtri_tiles(D, W.matrix, st, s);
    if (!err && alpha != 1.0)
        err = plasma_kern_geadd(0.0, B, alpha, B, st, s);
    const int steps = left ? B.mt : B.nt;
    for (int q = 0; q < steps && !err; q++) {
        const int k = forward ? q : steps - 1 - q;
        const int kn = left ? plasma_tile_mview(B, k) : plasma_tile_nview(B, k);
        const plasma_desc_t Wk = plasma_desc_view(W, 0, k * nb, kn, kn);
        /* rows (Left) / columns (Right) still to update */
        const int lo = forward ? k + 1 : 0, hi = forward ? steps : k;
        if (left) {
            const plasma_desc_t Bk = tview(B, k, k + 1, 0, B.nt);
            const plasma_desc_t Tk = plasma_desc_view(T, 0, 0, kn, B.n);
            err = plasma_kern_gemm(wop, PlasmaNoTrans, 1.0, Wk, Bk, 0.0, Tk,
                                   PlasmaGeneral, st, s);
            if (!err && lo < hi) {
                /* B(R, :) -= op(A)(R, k) T */
                const plasma_desc_t Ar = ta ? tview(A, k, k + 1, lo, hi)
                                            : tview(A, lo, hi, k, k + 1);
                err = plasma_kern_gemm(transa, PlasmaNoTrans, -1.0, Ar, Tk,
                                       1.0, tview(B, lo, hi, 0, B.nt),
                                       PlasmaGeneral, st, s);
            }
            if (!err)
                err = plasma_kern_lacpy(Tk, Bk, st, s);
        }
        else {
            const plasma_desc_t Bk = tview(B, 0, B.mt, k, k + 1);
            const plasma_desc_t Tk = plasma_desc_view(T, 0, 0, B.m, kn);
            err = plasma_kern_gemm(PlasmaNoTrans, wop, 1.0, Bk, Wk, 0.0, Tk,
                                   PlasmaGeneral, st, s);
            if (!err && lo < hi) {
                /* B(:, R) -= T op(A)(k, R) */
                const plasma_desc_t Ar = ta ? tview(A, lo, hi, k, k + 1)
                                            : tview(A, k, k + 1, lo, hi);
                err = plasma_kern_gemm(PlasmaNoTrans, transa, -1.0, Tk, Ar,
                                       1.0, tview(B, 0, B.mt, lo, hi),
                                       PlasmaGeneral, st, s);
            }
            if (!err)
                err = plasma_kern_lacpy(Tk, Bk, st, s);
        }
    }
    issue(err, sequence, request);
    /* the pool hands the blocks out again in stream order */
    plasma_desc_destroy(&D);
    plasma_desc_destroy(&W);
    plasma_desc_destroy(&T);
}

void plasma_omp_xtrsm(plasma_enum_t precision, plasma_enum_t side,
                      plasma_enum_t uplo, plasma_enum_t transa,
                      plasma_enum_t diag, double alpha, plasma_desc_t A,
                      plasma_desc_t B, plasma_sequence_t *sequence,
                      plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (side != PlasmaLeft && side != PlasmaRight) {
        plasma_error("illegal value of side");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (transa != PlasmaNoTrans && !is_trans(transa)) {
        plasma_error("illegal value of transa");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (diag != PlasmaUnit && diag != PlasmaNonUnit) {
        plasma_error("illegal value of diag");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (!desc_ok(A, "invalid A", sequence, request) ||
        !desc_ok(B, "invalid B", sequence, request) ||
        !common_checks(sequence, request))
        return;
    if (B.m == 0 || B.n == 0)
        return;
    if (!same_tiling(precision, A, B, sequence, request))
        return;
    const int an = side == PlasmaLeft ? B.m : B.n;
    if (A.m != A.n || A.m != an || A.mb > 512) {
        plasma_error("A must be square, of the order of B's side, with "
                     "tiles of at most 512");
        plasma_request_fail(sequence, request, PlasmaErrorIllegalValue);
        return;
    }
    if (sequence->status != PlasmaSuccess)
        return;
    pxtrsm(precision, side, uplo, transa, diag, alpha, A, B, sequence,
           request);
}

int plasma_xtrsm(plasma_enum_t precision, plasma_enum_t side,
                 plasma_enum_t uplo, plasma_enum_t transa, plasma_enum_t diag,
                 int m, int n, double alpha, void *pA, int lda, void *pB,
                 int ldb)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_fatal_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    if (side != PlasmaLeft && side != PlasmaRight) {
        plasma_error("illegal value of side");
        return -1;
    }
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        return -2;
    }
    if (transa != PlasmaNoTrans && !is_trans(transa)) {
        plasma_error("illegal value of transa");
        return -3;
    }
    if (diag != PlasmaUnit && diag != PlasmaNonUnit) {
        plasma_error("illegal value of diag");
        return -4;
    }
    if (m < 0) {
        plasma_error("illegal value of m");
        return -5;
    }
    if (n < 0) {
        plasma_error("illegal value of n");
        return -6;
    }
    const int an = side == PlasmaLeft ? m : n;
    if (lda < imax(1, an)) {
        plasma_error("illegal value of lda");
        return -8;
    }
    if (ldb < imax(1, m)) {
        plasma_error("illegal value of ldb");
        return -10;
    }
    if (m == 0 || n == 0)
        return PlasmaSuccess;

    const int nb = ctx->nb;
    if (nb > 512) {
        plasma_error("trsm needs a tile size of at most 512");
        return PlasmaErrorNotSupported;
    }
    plasma_desc_t A, B;
    int retval = plasma_desc_general_create(precision, nb, nb, an, an, 0, 0,
                                            an, an, &A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        return retval;
    }
    retval = plasma_desc_general_create(precision, nb, nb, m, n, 0, 0, m, n,
                                        &B);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_create() failed");
        plasma_desc_destroy(&A);
        return retval;
    }
    plasma_sequence_t sequence;
    plasma_request_t request;
    plasma_sequence_init(&sequence);
    plasma_request_init(&request);
    plasma_pxlapack2desc(PlasmaGeneral, pA, lda, A, &sequence, &request);
    plasma_pxlapack2desc(PlasmaGeneral, pB, ldb, B, &sequence, &request);
    plasma_omp_xtrsm(precision, side, uplo, transa, diag, alpha, A, B,
                     &sequence, &request);
    plasma_pxdesc2lapack(PlasmaGeneral, B, pB, ldb, &sequence, &request);
    int status = plasma_sequence_sync(&sequence);
    plasma_desc_destroy(&A);
    plasma_desc_destroy(&B);
    return status;
}

/***************************************************************************
 * Precision entry points.
 */
#define DEFINE_BLAS3(P, TYPE, PREC)                                           \
int plasma_##P##gemm(plasma_enum_t transa, plasma_enum_t transb, int m,      \
                     int n, int k, TYPE alpha, TYPE *pA, int lda, TYPE *pB,  \
                     int ldb, TYPE beta, TYPE *pC, int ldc)                  \
{                                                                             \
    return plasma_xgemm(PREC, transa, transb, m, n, k, alpha, pA, lda, pB,    \
                        ldb, beta, pC, ldc);                                  \
}                                                                             \
void plasma_omp_##P##gemm(plasma_enum_t transa, plasma_enum_t transb,        \
                          TYPE alpha, plasma_desc_t A, plasma_desc_t B,       \
                          TYPE beta, plasma_desc_t C,                         \
                          plasma_sequence_t *sequence,                        \
                          plasma_request_t *request)                          \
{                                                                             \
    plasma_omp_xgemm(PREC, transa, transb, alpha, A, B, beta, C, sequence,    \
                     request);                                                \
}                                                                             \
int plasma_##P##syrk(plasma_enum_t uplo, plasma_enum_t trans, int n, int k,  \
                     TYPE alpha, TYPE *pA, int lda, TYPE beta, TYPE *pC,     \
                     int ldc)                                                 \
{                                                                             \
    return plasma_xsyrk(PREC, uplo, trans, n, k, alpha, pA, lda, beta, pC,    \
                        ldc);                                                 \
}                                                                             \
void plasma_omp_##P##syrk(plasma_enum_t uplo, plasma_enum_t trans,           \
                          TYPE alpha, plasma_desc_t A, TYPE beta,             \
                          plasma_desc_t C, plasma_sequence_t *sequence,       \
                          plasma_request_t *request)                          \
{                                                                             \
    plasma_omp_xsyrk(PREC, uplo, trans, alpha, A, beta, C, sequence,          \
                     request);                                                \
}                                                                             \
int plasma_##P##trsm(plasma_enum_t side, plasma_enum_t uplo,                 \
                     plasma_enum_t transa, plasma_enum_t diag, int m, int n, \
                     TYPE alpha, TYPE *pA, int lda, TYPE *pB, int ldb)       \
{                                                                             \
    return plasma_xtrsm(PREC, side, uplo, transa, diag, m, n, alpha, pA,      \
                        lda, pB, ldb);                                        \
}                                                                             \
void plasma_omp_##P##trsm(plasma_enum_t side, plasma_enum_t uplo,            \
                          plasma_enum_t transa, plasma_enum_t diag,           \
                          TYPE alpha, plasma_desc_t A, plasma_desc_t B,       \
                          plasma_sequence_t *sequence,                        \
                          plasma_request_t *request)                          \
{                                                                             \
    plasma_omp_xtrsm(PREC, side, uplo, transa, diag, alpha, A, B, sequence,   \
                     request);                                                \
}

DEFINE_BLAS3(d, double, PlasmaRealDouble)
DEFINE_BLAS3(s, float, PlasmaRealFloat)
