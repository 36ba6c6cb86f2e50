/*
 * plasma_zc.h -- mixed precision (single complex factor, double complex
 * refinement) Hermitian solve and the conversions it uses.
 *
 * Signatures are the reference's (include/plasma_zc.h:33-37, :45-52,
 * :66-72, :82-88).  plasma_zcposv follows compute/zcposv.c:363-539 as
 * plasma_dsposv does: the Hermitian ||A||_inf (zlanhe), the column
 * max-norms max_i |Re x_ij| + |Im x_ij| (dzamax), the same stopping test,
 * at most 30 iterations, -31 and the double complex factorization when
 * refinement does not converge.  plasma_zlag2c returns 1 when the real or
 * the imaginary part of an entry overflows single precision.  Memory rules
 * and return codes as plasma_ds.h.
 */
#ifndef PLASMA_ZC_H
#define PLASMA_ZC_H

#include "plasma_async.h"
#include "plasma_descriptor.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference compute/zcposv.c:116 */
int plasma_zcposv(plasma_enum_t uplo, int n, int nrhs,
                  plasma_complex64_t *pA, int lda,
                  plasma_complex64_t *pB, int ldb,
                  plasma_complex64_t *pX, int ldx, int *iter);

/* reference compute/zcposv.c:363; W, Rnorm, Xnorm as plasma_omp_dsposv */
void plasma_omp_zcposv(plasma_enum_t uplo,
                       plasma_desc_t A, plasma_desc_t B, plasma_desc_t X,
                       plasma_desc_t As, plasma_desc_t Xs, plasma_desc_t R,
                       double *W, double *Rnorm, double *Xnorm, int *iter,
                       plasma_sequence_t *sequence, plasma_request_t *request);

/* reference compute/zlag2c.c:59 */
int plasma_zlag2c(int m, int n, plasma_complex64_t *pA, int lda,
                  plasma_complex32_t *pAs, int ldas);
/* reference compute/clag2z.c:58 */
int plasma_clag2z(int m, int n, plasma_complex32_t *pAs, int ldas,
                  plasma_complex64_t *pA, int lda);

/* reference include/plasma_zc.h:82-88 */
void plasma_omp_zlag2c(plasma_desc_t A, plasma_desc_t As,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_clag2z(plasma_desc_t As, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_ZC_H */
