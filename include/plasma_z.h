/*
 * plasma_z.h -- double complex (Hermitian) Cholesky solve path.
 *
 * Signatures are the reference's z forms (include/plasma_z.h:272-292,
 * :398-430, :631-646, :667-671).  A is Hermitian positive definite;
 * potrf computes A = L L^H (PlasmaLower) or A = U^H U (PlasmaUpper), using
 * the real part of A's diagonal, and leaves a real diagonal in the factor
 * (LAPACK zpotrf).  Elements are interleaved (re, im) pairs,
 * plasma_complex64_t.  Memory rules, return codes and the tile forms are
 * those of plasma_d.h.
 */
#ifndef PLASMA_Z_H
#define PLASMA_Z_H

#include "plasma_async.h"
#include "plasma_descriptor.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference compute/zpotrf.c:74 */
int plasma_zpotrf(plasma_enum_t uplo, int n, plasma_complex64_t *pA, int lda);
/* reference compute/zpotrs.c:77 */
int plasma_zpotrs(plasma_enum_t uplo, int n, int nrhs,
                  plasma_complex64_t *pA, int lda,
                  plasma_complex64_t *pB, int ldb);
/* reference compute/zposv.c:90 */
int plasma_zposv(plasma_enum_t uplo, int n, int nrhs,
                 plasma_complex64_t *pA, int lda,
                 plasma_complex64_t *pB, int ldb);

/* reference compute/zpotrf.c:206 */
void plasma_omp_zpotrf(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);
/* reference compute/zpotrs.c:226 */
void plasma_omp_zpotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request);
/* reference compute/zposv.c:244 */
void plasma_omp_zposv(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request);

/* reference compute/zge2desc.c:27, zdesc2ge.c, ztr2desc.c, zdesc2tr.c */
void plasma_omp_zge2desc(plasma_complex64_t *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_zdesc2ge(plasma_desc_t A, plasma_complex64_t *pA, int lda,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_ztr2desc(plasma_complex64_t *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_zdesc2tr(plasma_desc_t A, plasma_complex64_t *pA, int lda,
                         plasma_sequence_t *sequence, plasma_request_t *request);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_Z_H */
