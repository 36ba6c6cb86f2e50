/*
 * zchol.c -- tile Cholesky factorization and solve for the complex
 * precisions (PlasmaComplexDouble, PlasmaComplexFloat): the bodies of
 * plasma_pxpotrf / plasma_pxpotrs for complex descriptors.
 *
 * Reference: compute/pzpotrf.c:27-127 (A = L L^H lower, U^H U upper) and
 * compute/pztrsm.c:43-177 as called by zpotrs.c:263-275 (Left, NoTrans then
 * ConjTrans for lower; ConjTrans then NoTrans for upper).
 *
 * Per step k, on the library's main stream: the diagonal tile is factored
 * and inverted (W_kk) by one workgroup; the panel P(m) = L(m,k) W_kk^H is
 * one launch of 64 x 64 complex MFMA blocks into a panel buffer; one more
 * launch copies the panel back into L(m,k) and applies the herk / gemm
 * tasks of the step, L(m,n) -= P(m) P(n)^H.  Every step's update reads the
 * tiles in the order the reference's task DAG imposes on each tile, so each
 * tile receives its updates in increasing k.
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

void plasma_pzpotrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s = ctx->stream;
    const size_t es = plasma_element_size(A.precision);
    void *own_inv = NULL, *panel = NULL;
    if (inv == NULL &&
        (inv = own_inv = plasma_dev_alloc(plasma_inv_bytes(A))) == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    if (A.mt > 1) {
        panel = plasma_dev_alloc((size_t)(A.mt - 1) * A.mb * A.nb * es);
        if (panel == NULL) {
            plasma_dev_free(own_inv);
            plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
            return;
        }
    }
    for (int k = 0; k < A.mt; k++) {
        const double kv = plasma_tile_mview(A, k);
        const double rest = (double)A.m - (double)k * A.mb - kv;
        int slot;
        plasma_prof_begin(s, PlasmaProfPotrf, zpotrf_flops(kv), 0, &slot);
        int err = issue(plasma_kern_zpotrf_diag(uplo, A, k, inv, st, s),
                        sequence, request);
        plasma_prof_end(s, slot);
        if (err || k == A.mt - 1)
            break;
        plasma_prof_begin(s, PlasmaProfPanel, 8.0 * rest * kv * kv, 0, &slot);
        err = issue(plasma_kern_zpanel(uplo, A, k, inv, panel, st, s),
                    sequence, request);
        plasma_prof_end(s, slot);
        if (err)
            break;
        plasma_prof_begin(s, PlasmaProfUpdate, 4.0 * rest * (rest + 1) * kv,
                          0, &slot);
        err = issue(plasma_kern_zupdate(uplo, A, k, panel, st, s), sequence,
                    request);
        plasma_prof_end(s, slot);
        if (err)
            break;
    }
    /* the pool hands blocks out again in stream order */
    plasma_dev_free(panel);
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
