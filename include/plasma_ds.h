/*
 * plasma_ds.h -- mixed precision (FP32 factor, FP64 refinement) Cholesky
 * solve and the FP64 <-> FP32 conversions it uses.
 *
 * Signatures are the reference's ds forms (generated from
 * include/plasma_zc.h:33-37, :45-52, :66-72, :82-88).
 *
 * plasma_dsposv: factor A in FP32, refine X in FP64 until every column j
 * satisfies ||r_j||_max <= ||x_j||_max * ||A||_inf * eps64 * sqrt(n), at most
 * 30 iterations (reference compute/zcposv.c:363-539).  *iter returns 0 when
 * the FP32 solution already satisfies the test, k > 0 after k refinement
 * steps, and -31 when refinement did not converge and the FP64 factorization
 * was used instead.  pB is left unchanged; the solution goes to pX.  pA is
 * not modified (the FP64 fallback factors an internal copy).
 * Pointers may be host or device memory, as for plasma_dposv.
 */
#ifndef PLASMA_DS_H
#define PLASMA_DS_H

#include "plasma_async.h"
#include "plasma_descriptor.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference compute/zcposv.c:116 */
int plasma_dsposv(plasma_enum_t uplo, int n, int nrhs,
                  double *pA, int lda, double *pB, int ldb,
                  double *pX, int ldx, int *iter);

/* reference compute/zcposv.c:363.  A, B, X, R: FP64 general descriptors;
 * As, Xs: FP32 general descriptors; W, Rnorm, Xnorm: workspace the
 * reference needs on the host -- unused here (the reduction workspace and
 * the norms live on the device) but accepted for ABI compatibility and may
 * be NULL.  *iter is valid after plasma_sequence_sync(). */
void plasma_omp_dsposv(plasma_enum_t uplo,
                       plasma_desc_t A, plasma_desc_t B, plasma_desc_t X,
                       plasma_desc_t As, plasma_desc_t Xs, plasma_desc_t R,
                       double *W, double *Rnorm, double *Xnorm, int *iter,
                       plasma_sequence_t *sequence, plasma_request_t *request);

/* No reference counterpart (a mixed solver from a lower precision than
 * zcposv's): the FP64 solution of A X = B from a bf16-input factor.  A is
 * equilibrated symmetrically (S A S, S = diag(a_ii^{-1/2})), the scaled
 * matrix is factored in FP32 with its trailing updates on bf16 MFMA from
 * bf16 inputs (FP32 accumulation), and the solution is refined by GMRES-IR:
 * FP64 GMRES on M^{-1} A with the factor as preconditioner M, restarted
 * every 20 inner iterations, at most 10 refinement steps, with dsposv's
 * stopping test.  A factor that breaks down or a refinement that does not
 * converge falls back to the FP64 factorization.  *iter returns the total
 * inner GMRES iterations, or -1 after the fallback.  Arguments, return
 * codes and memory rules as plasma_dsposv (-10: iter is NULL). */
int plasma_dhposv(plasma_enum_t uplo, int n, int nrhs,
                  double *pA, int lda, double *pB, int ldb,
                  double *pX, int ldx, int *iter);

/* reference compute/zlag2c.c:59 -- returns 1 if an entry overflows FP32 */
int plasma_dlag2s(int m, int n, double *pA, int lda, float *pAs, int ldas);
/* reference compute/clag2z.c:58 */
int plasma_slag2d(int m, int n, float *pAs, int ldas, double *pA, int lda);

/* reference include/plasma_zc.h:82-88 */
void plasma_omp_dlag2s(plasma_desc_t A, plasma_desc_t As,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_slag2d(plasma_desc_t As, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_DS_H */
