/*
 * zchol.c -- tile Cholesky factorization and solve for the complex
 * precisions (PlasmaComplexDouble, PlasmaComplexFloat): the bodies of
 * plasma_pxpotrf / plasma_pxpotrs for complex descriptors.
 *
 * Reference: compute/pzpotrf.c:27-127 (A = L L^H lower, U^H U upper) and
 * compute/pztrsm.c:43-177 as called by zpotrs.c:263-275 (Left, NoTrans then
 * ConjTrans for lower; ConjTrans then NoTrans for upper).
 *
 * Per step k: the diagonal tile is factored and inverted (W_kk) by one
 * workgroup; the panel P(m) = L(m,k) W_kk^H is one launch of 64 x 64
 * complex MFMA blocks into a panel buffer; the herk / gemm tasks of the
 * step, L(m,n) -= P(m) P(n)^H, run as two launches (the next column on the
 * lookahead stream, the rest on the main stream, which also copies the
 * panel back into L(m,k)).  Each tile receives its updates in increasing
 * k, the order the reference's task DAG imposes.
 */
#include "internal.h"

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err != 0)
        plasma_request_fail(sequence, request, err);
    return err;
}

/* Real flops of complex products (reference test/flops.h: a complex
 * multiply-add is 8 real flops; zpotrf ~ 4 n^3 / 3). */
static double zpotrf_flops(double n)
{
    return 4.0 * n * n * n / 3.0;
}

/*
 * Lookahead of one column on two streams.  Per step k:
 *   panel stream:  wait bulk(k - 2);
 *                  narrow(k): column k's update by step k - 1's panels;
 *                  diagonal tile k (factor + W_kk); panel(k) into buffer
 *                  k % 2; record factor(k)
 *   main stream:   wait factor(k);
 *                  bulk(k): panels of step k copied into L(m,k), and the
 *                  update of columns k + 2 .. mt-1 by them; record bulk(k)
 * so the diagonal tile and panel of step k + 1 run beside the bulk update
 * of step k.  Races are excluded by construction: column k is written by
 * bulk(j) for j <= k - 2 (all waited for) and by narrow(k); buffer k % 2 is
 * rewritten by panel(k + 2) on the panel stream, after narrow(k + 1) read
 * it there and after the wait for bulk(k).
 */
#define ZEV_FACTOR 2
#define ZEV_BULK 3

void plasma_pzpotrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream, s1 = ctx->panel_stream;
    const size_t es = plasma_element_size(A.precision);
    const int mt = A.mt;
    void *own_inv = NULL, *panel[2] = {NULL, NULL};
    hipEvent_t ev[2 + ZEV_FACTOR + ZEV_BULK];
    int nev = 0, err = 0;
    if (inv == NULL &&
        (inv = own_inv = plasma_dev_alloc(plasma_inv_bytes(A))) == NULL)
        err = PlasmaErrorOutOfMemory;
    for (int i = 0; i < 2 && !err && mt > 1; i++)
        if ((panel[i] = plasma_dev_alloc((size_t)(mt - 1) * A.mb * A.nb *
                                         es)) == NULL)
            err = PlasmaErrorOutOfMemory;
    for (; nev < (int)(sizeof(ev) / sizeof(ev[0])) && !err; nev++)
        if (plasma_hip_check(hipEventCreateWithFlags(&ev[nev],
                                                     hipEventDisableTiming),
                             "hipEventCreate"))
            err = PlasmaErrorInternal;
    if (err) {
        plasma_request_fail(sequence, request, err);
        goto done;
    }
    hipEvent_t ev_fork = ev[0], ev_join = ev[1];
    hipEvent_t *ev_factor = ev + 2, *ev_bulk = ev + 2 + ZEV_FACTOR;

    hipEventRecord(ev_fork, s0);
    hipStreamWaitEvent(s1, ev_fork, 0);
    for (int k = 0; k < mt && !err; k++) {
        const double kv = plasma_tile_mview(A, k);
        const double rest = (double)A.m - (double)k * A.mb - kv;
        void *pk = panel[k % 2];
        int slot;
        /* panel stream */
        if (k >= 2)
            hipStreamWaitEvent(s1, ev_bulk[(k - 2) % ZEV_BULK], 0);
        if (k >= 1) {
            const double pv = plasma_tile_mview(A, k - 1);
            plasma_prof_begin(s1, PlasmaProfUpdate,
                              4.0 * (rest + kv) * (rest + kv + 1) * pv -
                                  4.0 * rest * (rest + 1) * pv,
                              0, &slot);
            err = issue(plasma_kern_zupdate(uplo, A, k - 1, k, k + 1, 0,
                                            panel[(k - 1) % 2], st, s1),
                        sequence, request);
            plasma_prof_end(s1, slot);
        }
        if (!err) {
            plasma_prof_begin(s1, PlasmaProfPotrf, zpotrf_flops(kv), 0,
                              &slot);
            err = issue(plasma_kern_zpotrf_diag(uplo, A, k, inv, st, s1),
                        sequence, request);
            plasma_prof_end(s1, slot);
        }
        if (!err && k < mt - 1) {
            plasma_prof_begin(s1, PlasmaProfPanel, 8.0 * rest * kv * kv, 0,
                              &slot);
            err = issue(plasma_kern_zpanel(uplo, A, k, inv, pk, st, s1),
                        sequence, request);
            plasma_prof_end(s1, slot);
        }
        hipEventRecord(ev_factor[k % ZEV_FACTOR], s1);
        /* main stream */
        hipStreamWaitEvent(s0, ev_factor[k % ZEV_FACTOR], 0);
        if (!err && k < mt - 1) {
            const double r2 = rest - plasma_tile_mview(A, k + 1);
            plasma_prof_begin(s0, PlasmaProfUpdate,
                              4.0 * r2 * (r2 + 1) * kv, 0, &slot);
            err = issue(plasma_kern_zupdate(uplo, A, k, k + 2, mt, 1, pk, st,
                                            s0),
                        sequence, request);
            plasma_prof_end(s0, slot);
        }
        hipEventRecord(ev_bulk[k % ZEV_BULK], s0);
    }
    hipEventRecord(ev_join, s1);
    hipStreamWaitEvent(s0, ev_join, 0);
done:
    for (int i = 0; i < nev; i++)
        hipEventDestroy(ev[i]);
    /* the pool hands blocks out again in stream order (main stream, which
     * has joined the panel stream) */
    plasma_dev_free(panel[0]);
    plasma_dev_free(panel[1]);
    plasma_dev_free(own_inv);
}

void plasma_pzpotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                    const void *inv, plasma_sequence_t *sequence,
                    plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s = ctx->stream;
    const size_t es = plasma_element_size(A.precision);
    void *own_inv = NULL;
    if (inv == NULL) {
        own_inv = plasma_dev_alloc(plasma_inv_bytes(A));
        if (own_inv == NULL) {
            plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
            return;
        }
        if (issue(plasma_kern_ztrtri_diag(uplo, A, own_inv, st, s), sequence,
                  request)) {
            plasma_dev_free(own_inv);
            return;
        }
        inv = own_inv;
    }
    void *S = plasma_dev_alloc((size_t)B.mb * B.nt * B.nb * es);
    if (S == NULL) {
        plasma_dev_free(own_inv);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    int slot;
    plasma_prof_begin(s, PlasmaProfSweep, 8.0 * 2.0 * B.m * B.m * B.n, 0,
                      &slot);
    int err = 0;
    /* L Y = B (for upper storage: U^H Y = B) */
    for (int k = 0; k < A.mt && !err; k++)
        err = issue(plasma_kern_zsweep_diag(0, B, k, inv, S, st, s),
                    sequence, request) ||
              issue(plasma_kern_zsweep_update(uplo, 0, A, B, k, S, st, s),
                    sequence, request);
    /* L^H X = Y (upper: U X = Y) */
    for (int k = A.mt - 1; k >= 0 && !err; k--)
        err = issue(plasma_kern_zsweep_diag(1, B, k, inv, S, st, s),
                    sequence, request) ||
              issue(plasma_kern_zsweep_update(uplo, 1, A, B, k, S, st, s),
                    sequence, request);
    plasma_prof_end(s, slot);
    plasma_dev_free(S);
    plasma_dev_free(own_inv);
}
