/*
 * plasma_batched.h -- many independent SPD systems of uneven order in one
 * launch.  No counterpart in the reference's interface (its drivers take
 * one matrix per call, include/plasma_z.h:272-292); the per-system results
 * follow LAPACK's potrf/posv: A is overwritten by its Cholesky factor in
 * the uplo triangle, B by the solution, info[i] = 0 or the order of the
 * first leading minor of system i that is not positive definite.
 *
 * Layout: system i, of order n[i] with 0 <= n[i] <= 512, is stored column
 * major at A + aoff[i] with leading dimension n[i]; its right-hand sides at
 * B + boff[i], n[i] x nrhs with leading dimension n[i] (offsets in
 * elements).  A and B must be device memory; n, aoff, boff and info may be
 * host or device arrays.  The call returns when the batch is done.
 *
 * Returns PlasmaSuccess, or -i when argument i is illegal (-3 also for an
 * order outside [0, 512]).
 */
#ifndef PLASMA_BATCHED_H
#define PLASMA_BATCHED_H

#include <stddef.h>
#include "plasma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int plasma_dposv_vbatched(plasma_enum_t uplo, int batch, const int *n,
                          double *A, const size_t *aoff, int nrhs, double *B,
                          const size_t *boff, int *info);
int plasma_sposv_vbatched(plasma_enum_t uplo, int batch, const int *n,
                          float *A, const size_t *aoff, int nrhs, float *B,
                          const size_t *boff, int *info);
int plasma_dpotrf_vbatched(plasma_enum_t uplo, int batch, const int *n,
                           double *A, const size_t *aoff, int *info);
int plasma_spotrf_vbatched(plasma_enum_t uplo, int batch, const int *n,
                           float *A, const size_t *aoff, int *info);

/*
 * plasma_dposv_vbatched without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results, info values and error codes as
 * plasma_dposv_vbatched; in addition batch > PLASMA_VBATCH_NOALLOC_MAX
 * returns -2.  Its scratch (the persistent grid's diagonal-block inverses,
 * the order of the systems, the ticket, and staged copies of host n / aoff
 * / boff / info) is one fixed device array per GPU of about 36 MB, reserved
 * when the library loads and kept until the process ends, so it calls no
 * hipMalloc, hipFree or allocating library routine.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several
 * host threads) race; calls in sequence are safe.  The call returns when
 * the batch is done.
 */
#define PLASMA_VBATCH_NOALLOC_MAX  65536
#define PLASMA_VBATCH_NOALLOC_GRID 256
int plasma_dposv_vbatched_noalloc(plasma_enum_t uplo, int batch,
                                  const int *n, double *A,
                                  const size_t *aoff, int nrhs, double *B,
                                  const size_t *boff, int *info);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_BATCHED_H */
