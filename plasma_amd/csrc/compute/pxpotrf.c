/*
 * pxpotrf.c -- tile Cholesky and its triangular sweeps as lookahead-
 * scheduled, stream-ordered launch sequences.
 *
 * The reference inserts one OpenMP task per tile operation and lets the
 * runtime order them by their depend clauses (compute/pzpotrf.c:37-80 for
 * lower, :84-126 for upper; 341 376 gemm tasks at n = 32768, nb = 256).
 * Here every step k issues one launch per task class covering all tiles of
 * that step:
 *     potrf(k)            diagonal tile, plus W_kk = L(k,k)^{-1}
 *     trsm(k)             panel tiles m > k: A(m,k) W_kk^T (MFMA)
 *     update(k, n0, n1)   trailing tiles of columns n0..n1 (MFMA)
 * With lookahead depth >= 1 the update of step k is split: column k+1 is
 * updated first on the main stream, the panel stream then factors column
 * k+1 (potrf + trsm) while the main stream updates the remaining columns.
 * The panel solve writes into a panel buffer that the updates read; the
 * copy into A(:, k) rides along with the bulk update of step k, off the
 * critical path.
 * Events carry the two dependencies the depend clauses expressed:
 *     panel(k+1) after update(k, k+1)           (ev_col)
 *     update(k+1, ...) after panel(k+1)         (ev_panel)
 * so the critical path potrf -> trsm -> next column overlaps the bulk
 * update.  The streams are joined again before returning.
 */
#include "internal.h"

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err != 0)
        plasma_request_fail(sequence, request, err);
    return err;
}

size_t plasma_inv_bytes(plasma_desc_t A)
{
    return (size_t)A.mt * A.mb * A.mb * plasma_element_size(A.precision);
}

/* Algorithmic flops of the launches (reference test/flops.h conventions:
 * potrf n^3/3 + n^2/2 + n/6, trsm m n^2, syrk k n (n + 1), gemm 2 m n k). */
static double potrf_flops(double n)
{
    return n * n * n / 3.0 + n * n / 2.0 + n / 6.0;
}

static double panel_flops(plasma_desc_t A, int k)
{
    double f = 0, nb = A.nb;
    for (int m = k + 1; m < A.mt; m++)
        f += plasma_tile_mview(A, m) * nb * nb;
    return f;
}

static double update_flops(plasma_desc_t A, int n0, int n1)
{
    double f = 0, nb = A.nb;
    for (int n = n0; n < n1 && n < A.mt; n++) {
        double nv = plasma_tile_mview(A, n);
        f += nb * nv * (nv + 1);                           /* syrk (n, n) */
        for (int m = n + 1; m < A.mt; m++)
            f += 2.0 * plasma_tile_mview(A, m) * nv * nb;  /* gemm */
    }
    return f;
}

static int k_potrf(plasma_enum_t uplo, plasma_desc_t A, int k, void *inv,
                   int *st, hipStream_t s)
{
    int slot;
    plasma_prof_begin(s, PlasmaProfPotrf, potrf_flops(plasma_tile_mview(A, k)),
                      0, &slot);
    int err = plasma_kern_potrf_diag(uplo, A, k, inv, st, s);
    plasma_prof_end(s, slot);
    return err;
}

static int k_panel(plasma_enum_t uplo, plasma_desc_t A, int k, void *inv,
                   void *panel, int *st, hipStream_t s)
{
    int slot;
    plasma_prof_begin(s, PlasmaProfPanel, panel_flops(A, k), 0, &slot);
    int err = plasma_kern_trsm_panel(uplo, A, k, inv, panel, st, s);
    plasma_prof_end(s, slot);
    return err;
}

static int k_update(plasma_enum_t uplo, plasma_desc_t A, int k, int n0,
                    int n1, const void *panel, int copy_back, int *st,
                    hipStream_t s)
{
    int slot;
    plasma_context_t *ctx = plasma_context_self();
    plasma_prof_begin(s, PlasmaProfUpdate,
                      ctx->profile ? update_flops(A, n0, n1) : 0, 0, &slot);
    int err = plasma_kern_trailing_update(uplo, A, k, n0, n1, panel,
                                          copy_back, st, s);
    plasma_prof_end(s, slot);
    return err;
}

void plasma_pxpotrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream, s1 = ctx->panel_stream;
    const int mt = A.mt;
    size_t es = plasma_element_size(A.precision);
    const int lookahead = ctx->lookahead > 0 && mt >= 3;

    void *own_inv = NULL;
    if (inv == NULL) {
        inv = own_inv = plasma_dev_alloc(plasma_inv_bytes(A));
        if (inv == NULL) {
            plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
            return;
        }
    }
    /* panel buffers: the solved panel of step k lives in panel[k % 2] until
     * the rest of step k's update has copied it into A(:, k) */
    const size_t pbytes = (size_t)imax(mt - 1, 1) * A.mb * A.nb * es;
    void *panel[2] = {NULL, NULL};
    panel[0] = plasma_dev_alloc(pbytes);
    panel[1] = lookahead ? plasma_dev_alloc(pbytes) : panel[0];
    if (panel[0] == NULL || panel[1] == NULL) {
        if (panel[1] != panel[0])
            plasma_dev_free(panel[1]);
        plasma_dev_free(panel[0]);
        plasma_dev_free(own_inv);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }

    if (!lookahead) {
        for (int k = 0; k < mt; k++) {
            if (issue(k_potrf(uplo, A, k, inv, st, s0), sequence, request) ||
                issue(k_panel(uplo, A, k, inv, panel[0], st, s0), sequence,
                      request) ||
                issue(k_update(uplo, A, k, k + 1, mt, panel[0], 1, st, s0),
                      sequence, request))
                break;
        }
        plasma_dev_free(panel[0]);
        plasma_dev_free(own_inv);
        return;
    }

    hipEvent_t ev_fork, ev_col, ev_panel;
    if (plasma_hip_check(hipEventCreateWithFlags(&ev_fork,
                                                 hipEventDisableTiming),
                         "hipEventCreate") ||
        plasma_hip_check(hipEventCreateWithFlags(&ev_col,
                                                 hipEventDisableTiming),
                         "hipEventCreate") ||
        plasma_hip_check(hipEventCreateWithFlags(&ev_panel,
                                                 hipEventDisableTiming),
                         "hipEventCreate")) {
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        plasma_dev_free(panel[1]);
        plasma_dev_free(panel[0]);
        plasma_dev_free(own_inv);
        return;
    }

    /* fork: the panel stream starts after everything queued so far */
    hipEventRecord(ev_fork, s0);
    hipStreamWaitEvent(s1, ev_fork, 0);

    int err = issue(k_potrf(uplo, A, 0, inv, st, s1), sequence, request);
    if (!err)
        err = issue(k_panel(uplo, A, 0, inv, panel[0], st, s1), sequence,
                    request);
    hipEventRecord(ev_panel, s1);

    for (int k = 0; k + 1 < mt && !err; k++) {
        void *pk = panel[k & 1];
        /* main stream: wait for panel k, update column k+1 */
        hipStreamWaitEvent(s0, ev_panel, 0);
        err = issue(k_update(uplo, A, k, k + 1, k + 2, pk, 0, st, s0),
                    sequence, request);
        hipEventRecord(ev_col, s0);
        /* panel stream: factor column k+1 into the other panel buffer,
         * whose last reader (step k-1's update and copy-back) precedes
         * ev_col on the main stream */
        hipStreamWaitEvent(s1, ev_col, 0);
        if (!err)
            err = issue(k_potrf(uplo, A, k + 1, inv, st, s1), sequence,
                        request);
        if (!err)
            err = issue(k_panel(uplo, A, k + 1, inv, panel[(k + 1) & 1], st,
                                s1), sequence, request);
        hipEventRecord(ev_panel, s1);
        /* main stream: the rest of the trailing matrix, and panel k back
         * into A(:, k) */
        if (!err)
            err = issue(k_update(uplo, A, k, k + 2, mt, pk, 1, st, s0),
                        sequence, request);
    }
    /* join */
    hipEventRecord(ev_panel, s1);
    hipStreamWaitEvent(s0, ev_panel, 0);
    hipEventDestroy(ev_fork);
    hipEventDestroy(ev_col);
    hipEventDestroy(ev_panel);
    plasma_dev_free(panel[1]);
    plasma_dev_free(panel[0]);
    plasma_dev_free(own_inv);
}

/*
 * Triangular sweeps (reference compute/pztrsm.c, Left side).  In the
 * lower-equivalent convention the four (uplo, trans) cases used by potrs
 * reduce to a forward sweep with L or a backward sweep with L^T.
 */
void plasma_pxtrsm_left(plasma_enum_t uplo, plasma_enum_t trans,
                        plasma_desc_t A, plasma_desc_t B, const void *inv,
                        plasma_sequence_t *sequence,
                        plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    /* lower+NoTrans and upper+Trans solve with L; the others with L^T */
    const int backward = (uplo == PlasmaLower) ? (trans != PlasmaNoTrans)
                                               : (trans == PlasmaNoTrans);
    const int mt = B.mt;
    size_t es = plasma_element_size(B.precision);
    void *scratch = plasma_dev_alloc((size_t)A.mb * B.nt * B.nb * es);
    if (scratch == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    for (int i = 0; i < mt; i++) {
        const int k = backward ? mt - 1 - i : i;
        int slot;
        double nbk = plasma_tile_mview(B, k);
        int nupd = backward ? k : mt - 1 - k;
        plasma_prof_begin(s0, PlasmaProfSweep,
                          nbk * nbk * B.n + 2.0 * nbk * nupd * B.mb * B.n, 0,
                          &slot);
        int err = issue(plasma_kern_trsm_diag_left(uplo, backward, A, B, k,
                                                   inv, scratch, st, s0),
                        sequence, request);
        if (!err)
            err = issue(plasma_kern_sweep_update(uplo, backward, A, B, k,
                                                 scratch, st, s0),
                        sequence, request);
        plasma_prof_end(s0, slot);
        if (err)
            break;
    }
    plasma_dev_free(scratch);
}

void plasma_pxpotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                    const void *inv, plasma_sequence_t *sequence,
                    plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    void *own_inv = NULL;
    if (inv == NULL) {
        plasma_context_t *ctx = plasma_context_self();
        int *st = plasma_sequence_dstatus(sequence, request);
        own_inv = plasma_dev_alloc(plasma_inv_bytes(A));
        if (own_inv == NULL) {
            plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
            return;
        }
        if (issue(plasma_kern_trtri_diag(uplo, A, own_inv, st, ctx->stream),
                  sequence, request)) {
            plasma_dev_free(own_inv);
            return;
        }
        inv = own_inv;
    }
    /* reference compute/zpotrs.c:270-283 */
    plasma_pxtrsm_left(uplo, uplo == PlasmaUpper ? PlasmaConjTrans
                                                 : PlasmaNoTrans,
                       A, B, inv, sequence, request);
    plasma_pxtrsm_left(uplo, uplo == PlasmaUpper ? PlasmaNoTrans
                                                 : PlasmaConjTrans,
                       A, B, inv, sequence, request);
    plasma_dev_free(own_inv);
}

/*
 * Repeated solves with one factor (the refinement loop of dsposv).  With
 * L = M D, M(m,k) = L(m,k) W_kk (m > k), M(k,k) = I and D = diag(L_kk):
 *     L Y = B      <=>  Z = M^{-1} B,  Y = D^{-1} Z
 *     L^T X = Y    <=>  M^T X = D^{-T} Y = W^T W Z
 * so each sweep step is one launch (B(k,:) is final when step k runs)
 * instead of a diagonal solve plus an update; the two diagonal products
 * run over all block rows at once.  The reference's pztrsm sweeps
 * (compute/pztrsm.c:113-177) compute the same solution.
 */
size_t plasma_mpanel_bytes(plasma_desc_t A)
{
    return (size_t)A.mt * (A.mt - 1) / 2 * A.mb * A.nb *
           plasma_element_size(A.precision);
}

void plasma_pxscale_panels(plasma_enum_t uplo, plasma_desc_t A,
                           const void *inv, void *M,
                           plasma_sequence_t *sequence,
                           plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    issue(plasma_kern_scale_panels(uplo, A, inv, M, st, ctx->stream),
          sequence, request);
}

void plasma_pxpotrs_m(plasma_desc_t A, plasma_desc_t B, const void *inv,
                      const void *M, plasma_sequence_t *sequence,
                      plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    const int mt = B.mt;
    size_t es = plasma_element_size(B.precision);
    void *scratch = plasma_dev_alloc((size_t)mt * A.mb * B.nt * B.nb * es);
    if (scratch == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    int err = 0, slot;
    const double sweep_flops = 2.0 * B.m * B.m * B.n;  /* both sweeps */
    plasma_prof_begin(s0, PlasmaProfSweep, sweep_flops, 0, &slot);
    for (int k = 0; k + 1 < mt && !err; k++)
        err = issue(plasma_kern_msweep_update(0, A, B, k, M, st, s0),
                    sequence, request);
    if (!err)
        err = issue(plasma_kern_diag_normal(A, B, inv, scratch, st, s0),
                    sequence, request);
    for (int k = mt - 1; k > 0 && !err; k--)
        err = issue(plasma_kern_msweep_update(1, A, B, k, M, st, s0),
                    sequence, request);
    plasma_prof_end(s0, slot);
    plasma_dev_free(scratch);
}
