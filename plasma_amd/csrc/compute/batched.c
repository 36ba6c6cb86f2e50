/*
 * batched.c -- SPD solves of many independent systems of uneven order in
 * one persistent launch (plasma_{d,s}posv_vbatched, plasma_{d,s}potrf_
 * vbatched).  No counterpart in the reference, which factors one matrix
 * per call (compute/zposv.c); each system follows the same algorithm as
 * the reference's tile factorization of a single tile (core_zpotrf.c) and
 * its triangular solves, with LAPACK's info per system.
 *
 * Layout: system i of order n[i] (0 <= n[i] <= 512) is stored column major
 * at A + aoff[i] with leading dimension n[i]; its right-hand sides at
 * B + boff[i], n[i] x nrhs, leading dimension n[i].  A and B are device
 * memory; n, aoff, boff and info may be host or device arrays.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define VBATCH_NMAX 512

typedef struct {
    int n, i;
} order_t;

static int by_n_desc(const void *a, const void *b)
{
    const order_t *x = a, *y = b;
    if (x->n != y->n)
        return y->n - x->n;
    return x->i - y->i;
}

/* Device copy of a host or device array (owned copy in *dev_out).  The
 * copy goes on the context stream: pool memory is handed out in that
 * stream's order, and a block returned by an earlier call may still be read
 * by its queued kernels, so a copy on another stream (or the null stream,
 * which the library's non-blocking streams do not wait for) could overwrite
 * it under them. */
static int stage_in(const void *arr, size_t bytes, void **dev_out,
                    int *owned, hipStream_t s)
{
    if (plasma_is_device_ptr(arr)) {
        *dev_out = (void *)arr;
        *owned = 0;
        return 0;
    }
    *dev_out = plasma_dev_alloc(bytes);
    if (*dev_out == NULL)
        return PlasmaErrorOutOfMemory;
    *owned = 1;
    return plasma_hip_check(hipMemcpyAsync(*dev_out, arr, bytes,
                                           hipMemcpyHostToDevice, s),
                            "hipMemcpyAsync")
               ? PlasmaErrorInternal : 0;
}

/* Argument checks of the batched drivers, in the order of their codes:
 * returns -i for an illegal argument i, 1 for the quick return (batch 0),
 * 0 to proceed. */
static int vbatched_args(plasma_enum_t uplo, int batch, const int *n,
                         const void *A, const size_t *aoff, int nrhs,
                         const void *B, const size_t *boff, const int *info)
{
    if (uplo != PlasmaUpper && uplo != PlasmaLower) {
        plasma_error("illegal value of uplo");
        return -1;
    }
    if (batch < 0) {
        plasma_error("illegal value of batch");
        return -2;
    }
    if (batch == 0)
        return 1;
    if (n == NULL) {
        plasma_error("NULL n");
        return -3;
    }
    if (A == NULL || !plasma_is_device_ptr(A)) {
        plasma_error("A must be device memory");
        return -4;
    }
    if (aoff == NULL) {
        plasma_error("NULL aoff");
        return -5;
    }
    if (nrhs < 0) {
        plasma_error("illegal value of nrhs");
        return -6;
    }
    if (nrhs > 0 && (B == NULL || !plasma_is_device_ptr(B))) {
        plasma_error("B must be device memory");
        return -7;
    }
    if (nrhs > 0 && boff == NULL) {
        plasma_error("NULL boff");
        return -8;
    }
    if (info == NULL) {
        plasma_error("NULL info");
        return -9;
    }
    return 0;
}

/* Workgroups of the persistent grid: as many as the LDS lets the CUs hold,
 * at most one per system. */
static int vbatched_grid(plasma_context_t *ctx, plasma_enum_t precision,
                         int batch)
{
    const size_t lds = plasma_kern_vbatched_lds(precision);
    int per_cu = (int)((160 * 1024) / lds);
    if (per_cu < 1)
        per_cu = 1;
    int grid = ctx->num_cus * per_cu;
    return grid > batch ? batch : grid;
}

static int vbatched(plasma_enum_t precision, plasma_enum_t uplo, int batch,
                    const int *n, void *A, const size_t *aoff, int nrhs,
                    void *B, const size_t *boff, int *info)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    const int args = vbatched_args(uplo, batch, n, A, aoff, nrhs, B, boff,
                                   info);
    if (args != 0)
        return args < 0 ? args : PlasmaSuccess;

    /* host view of n: the argument check and the order */
    const size_t nbytes = (size_t)batch * sizeof(int);
    const size_t obytes = (size_t)batch * sizeof(size_t);
    int *hn = malloc(nbytes);
    order_t *ord = malloc((size_t)batch * sizeof(order_t));
    int *hord = malloc(nbytes);
    int retval = PlasmaSuccess;
    void *dn = NULL, *doff = NULL, *dboff = NULL, *dinfo = NULL;
    int own_n = 0, own_off = 0, own_boff = 0, own_info = 0;
    void *W = NULL, *dord = NULL, *dticket = NULL;
    if (hn == NULL || ord == NULL || hord == NULL) {
        retval = PlasmaErrorOutOfMemory;
        goto done;
    }
    /* a caller's device array may have been written by work queued on the
     * library's streams (non-blocking, so the copy below would not wait for
     * it) */
    if (plasma_is_device_ptr(n)) {
        hipStreamSynchronize(ctx->panel_stream);
        hipStreamSynchronize(ctx->stream);
    }
    if (plasma_hip_check(hipMemcpy(hn, n, nbytes, hipMemcpyDefault),
                         "hipMemcpy(n)")) {
        retval = PlasmaErrorInternal;
        goto done;
    }
    for (int i = 0; i < batch; i++) {
        if (hn[i] < 0 || hn[i] > VBATCH_NMAX) {
            plasma_error("illegal order of a system");
            retval = -3;
            goto done;
        }
        ord[i].n = hn[i];
        ord[i].i = i;
    }
    qsort(ord, batch, sizeof(order_t), by_n_desc);
    for (int i = 0; i < batch; i++)
        hord[i] = ord[i].i;

    /* one resident grid, each workgroup with its own slab of W */
    const int grid = vbatched_grid(ctx, precision, batch);
    W = plasma_dev_alloc((size_t)grid * plasma_kern_vbatched_wslab(precision));
    dord = plasma_dev_alloc(nbytes);
    dticket = plasma_dev_alloc(sizeof(int));
    if (W == NULL || dord == NULL || dticket == NULL) {
        retval = PlasmaErrorOutOfMemory;
        goto done;
    }
    if ((retval = stage_in(n, nbytes, &dn, &own_n, ctx->stream)) != 0 ||
        (retval = stage_in(aoff, obytes, &doff, &own_off, ctx->stream)) != 0 ||
        (nrhs > 0 &&
         (retval = stage_in(boff, obytes, &dboff, &own_boff, ctx->stream)) !=
             0))
        goto done;
    own_info = !plasma_is_device_ptr(info);
    if (own_info) {
        dinfo = plasma_dev_alloc(nbytes);
        if (dinfo == NULL) {
            retval = PlasmaErrorOutOfMemory;
            goto done;
        }
    }
    else {
        dinfo = info;
    }
    hipStream_t s = ctx->stream;
    if (plasma_hip_check(hipMemcpyAsync(dord, hord, nbytes,
                                        hipMemcpyHostToDevice, s),
                         "hipMemcpy(order)") ||
        plasma_kern_fill_int(dticket, 0, 1, s) != 0) {
        retval = PlasmaErrorInternal;
        goto done;
    }
    if (plasma_kern_vbatched_posv(precision, uplo == PlasmaUpper, batch, dn,
                                  A, doff, W, nrhs, B, dboff, dinfo, dord,
                                  dticket, grid, s) != 0) {
        retval = PlasmaErrorInternal;
        goto done;
    }
    if (own_info &&
        plasma_hip_check(hipMemcpyAsync(info, dinfo, nbytes,
                                        hipMemcpyDeviceToHost, s),
                         "hipMemcpy(info)")) {
        retval = PlasmaErrorInternal;
        goto done;
    }
    if (plasma_hip_check(hipStreamSynchronize(s), "hipStreamSynchronize"))
        retval = PlasmaErrorInternal;

done:
    /* the host arrays below are freed only after the copies that read them
     * (also on the error paths) */
    hipStreamSynchronize(ctx->stream);
    if (own_n) plasma_dev_free(dn);
    if (own_off) plasma_dev_free(doff);
    if (own_boff) plasma_dev_free(dboff);
    if (own_info) plasma_dev_free(dinfo);
    plasma_dev_free(dticket);
    plasma_dev_free(dord);
    plasma_dev_free(W);
    free(hord);
    free(ord);
    free(hn);
    return retval;
}

int plasma_dposv_vbatched(plasma_enum_t uplo, int batch, const int *n,
                          double *A, const size_t *aoff, int nrhs, double *B,
                          const size_t *boff, int *info)
{
    return vbatched(PlasmaRealDouble, uplo, batch, n, A, aoff, nrhs, B, boff,
                    info);
}

int plasma_sposv_vbatched(plasma_enum_t uplo, int batch, const int *n,
                          float *A, const size_t *aoff, int nrhs, float *B,
                          const size_t *boff, int *info)
{
    return vbatched(PlasmaRealFloat, uplo, batch, n, A, aoff, nrhs, B, boff,
                    info);
}

int plasma_dpotrf_vbatched(plasma_enum_t uplo, int batch, const int *n,
                           double *A, const size_t *aoff, int *info)
{
    return vbatched(PlasmaRealDouble, uplo, batch, n, A, aoff, 0, NULL, NULL,
                    info);
}

int plasma_spotrf_vbatched(plasma_enum_t uplo, int batch, const int *n,
                           float *A, const size_t *aoff, int *info)
{
    return vbatched(PlasmaRealFloat, uplo, batch, n, A, aoff, 0, NULL, NULL,
                    info);
}

/* The allocation-free variant: every device buffer beyond the caller's
 * inputs and outputs is a region of the per-device scratch array of
 * kernels/noalloc.hip (its address looked up once per device), and no host
 * buffer is allocated either.  Host metadata arrays are copied into the
 * scratch; the systems are ordered on the device. */
int plasma_dposv_vbatched_noalloc(plasma_enum_t uplo, int batch,
                                  const int *n, double *A,
                                  const size_t *aoff, int nrhs, double *B,
                                  const size_t *boff, int *info)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    const int args = vbatched_args(uplo, batch, n, A, aoff, nrhs, B, boff,
                                   info);
    if (args != 0)
        return args < 0 ? args : PlasmaSuccess;
    if (batch > PLASMA_VBATCH_NOALLOC_MAX) {
        plasma_error("batch above PLASMA_VBATCH_NOALLOC_MAX, the number of "
                     "systems the allocation-free scratch holds");
        return -2;
    }
    plasma_noalloc_t r;
    if (plasma_noalloc_setup(ctx->device) != 0 ||
        plasma_noalloc_regions(ctx->device, &r) != 0 ||
        plasma_kern_vbatched_wslab(PlasmaRealDouble) > r.wslab_bytes) {
        plasma_error("allocation-free scratch unavailable");
        return PlasmaErrorInternal;
    }
    hipStream_t s = ctx->stream;
    const size_t nbytes = (size_t)batch * sizeof(int);
    const size_t obytes = (size_t)batch * sizeof(size_t);
    const int host_n = !plasma_is_device_ptr(n);
    const int *dn = n;
    if (host_n) {
        /* checked here, like plasma_dposv_vbatched, before any launch */
        for (int i = 0; i < batch; i++)
            if (n[i] < 0 || n[i] > VBATCH_NMAX) {
                plasma_error("illegal order of a system");
                return -3;
            }
        if (plasma_hip_check(hipMemcpyAsync(r.n, n, nbytes,
                                            hipMemcpyHostToDevice, s),
                             "hipMemcpyAsync(n)"))
            return PlasmaErrorInternal;
        dn = r.n;
    }
    const size_t *daoff = aoff, *dboff = boff;
    if (!plasma_is_device_ptr(aoff)) {
        if (plasma_hip_check(hipMemcpyAsync(r.aoff, aoff, obytes,
                                            hipMemcpyHostToDevice, s),
                             "hipMemcpyAsync(aoff)"))
            return PlasmaErrorInternal;
        daoff = r.aoff;
    }
    if (nrhs > 0 && !plasma_is_device_ptr(boff)) {
        if (plasma_hip_check(hipMemcpyAsync(r.boff, boff, obytes,
                                            hipMemcpyHostToDevice, s),
                             "hipMemcpyAsync(boff)"))
            return PlasmaErrorInternal;
        dboff = r.boff;
    }
    const int host_info = !plasma_is_device_ptr(info);
    int *dinfo = host_info ? r.info : info;
    if (plasma_kern_vbatched_order(dn, batch, r.order, r.ctrl, s) != 0)
        return PlasmaErrorInternal;
    if (!host_n) {
        /* orders in device memory: checked by the order kernel */
        int ctrl[2] = {0, 0};
        if (plasma_hip_check(hipMemcpyAsync(ctrl, r.ctrl, sizeof(ctrl),
                                            hipMemcpyDeviceToHost, s),
                             "hipMemcpyAsync(flag)") ||
            plasma_hip_check(hipStreamSynchronize(s), "hipStreamSynchronize"))
            return PlasmaErrorInternal;
        if (ctrl[1] != 0) {
            plasma_error("illegal order of a system");
            return -3;
        }
    }
    int grid = vbatched_grid(ctx, PlasmaRealDouble, batch);
    if (grid > r.grid_max)
        grid = r.grid_max;
    if (plasma_kern_vbatched_posv(PlasmaRealDouble, uplo == PlasmaUpper,
                                  batch, dn, A, daoff, r.W, nrhs, B, dboff,
                                  dinfo, r.order, r.ctrl, grid, s) != 0)
        return PlasmaErrorInternal;
    if (host_info &&
        plasma_hip_check(hipMemcpyAsync(info, r.info, nbytes,
                                        hipMemcpyDeviceToHost, s),
                         "hipMemcpyAsync(info)"))
        return PlasmaErrorInternal;
    return plasma_hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")
               ? PlasmaErrorInternal : PlasmaSuccess;
}
