/*
 * plasma_s.h -- single precision Cholesky solve path.
 *
 * Signatures are the reference's s-precision forms (generated from
 * include/plasma_z.h:272-292, :398-430, :631-646, :667-671).
 *
 * LAPACK-layout routines (plasma_spotrf, plasma_spotrs, plasma_sposv) take
 * column-major arrays that may live in host memory or in device memory of
 * the library's GPU; the library detects which.  Host arrays are staged
 * through HBM; device arrays are converted in place on the GPU.  The call
 * returns when all work has finished, like the reference's implicit
 * barrier.
 *
 * Tile routines (plasma_omp_*) take descriptors whose storage is device
 * memory and enqueue their work on the library's streams; completion and
 * the final sequence->status come from plasma_sequence_sync().  The
 * pA arguments of the layout-translation routines may be host or device
 * pointers; for host pointers those two routines complete before they
 * return.
 *
 * Return codes: 0 success; -i: the i-th argument had an illegal value;
 * i > 0: the leading minor of order i is not positive definite (global,
 * 1-based, as LAPACK dpotrf); negative PlasmaError* codes for library
 * errors.
 */
#ifndef PLASMA_S_H
#define PLASMA_S_H

#include "plasma_async.h"
#include "plasma_descriptor.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference compute/zpotrf.c:74 */
int plasma_spotrf(plasma_enum_t uplo, int n, float *pA, int lda);
/* reference compute/zpotrs.c:77 */
int plasma_spotrs(plasma_enum_t uplo, int n, int nrhs,
                  float *pA, int lda, float *pB, int ldb);
/* reference compute/zposv.c:90 */
int plasma_sposv(plasma_enum_t uplo, int n, int nrhs,
                 float *pA, int lda, float *pB, int ldb);

/* reference compute/zpotrf.c:206 */
void plasma_omp_spotrf(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);
/* reference compute/zpotrs.c:226 */
void plasma_omp_spotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request);
/* reference compute/zposv.c:244 */
void plasma_omp_sposv(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request);

/* reference compute/zge2desc.c:27, zdesc2ge.c, ztr2desc.c, zdesc2tr.c */
void plasma_omp_sge2desc(float *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_sdesc2ge(plasma_desc_t A, float *pA, int lda,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_str2desc(float *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_sdesc2tr(plasma_desc_t A, float *pA, int lda,
                         plasma_sequence_t *sequence, plasma_request_t *request);

/* Inverse of an SPD matrix (reference compute/zpotri.c:65, compute/
 * zpoinv.c:67), single precision forms of plasma_d.h's. */
int plasma_spotri(plasma_enum_t uplo, int n, float *pA, int lda);
int plasma_spoinv(plasma_enum_t uplo, int n, float *pA, int lda);
void plasma_omp_spotri(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_spoinv(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);

/* Norms (reference compute/zlansy.c:81, :240; compute/zlange.c:76, :233;
 * compute/dzamax.c:21, :113), single precision forms of plasma_d.h's. */
float plasma_slansy(plasma_enum_t norm, plasma_enum_t uplo, int n,
                    float *pA, int lda);
float plasma_slange(plasma_enum_t norm, int m, int n, float *pA, int lda);
int plasma_samax(plasma_enum_t colrow, int m, int n, float *pA, int lda,
                 float *values);
void plasma_omp_slansy(plasma_enum_t norm, plasma_enum_t uplo,
                       plasma_desc_t A, double *work, float *value,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_slange(plasma_enum_t norm, plasma_desc_t A, double *work,
                       float *value, plasma_sequence_t *sequence,
                       plasma_request_t *request);
void plasma_omp_samax(plasma_enum_t colrow, plasma_desc_t A, double *work,
                      float *values, plasma_sequence_t *sequence,
                      plasma_request_t *request);

/*
 * BLAS-3 (reference compute/zgemm.c:105, :315; compute/zsyrk.c:89, :263;
 * compute/ztrsm.c:104, :290).  C = alpha op(A) op(B) + beta C;
 * C = alpha A A^T + beta C (NoTrans) or alpha A^T A + beta C (Trans) in the
 * uplo triangle of C; B = alpha op(A)^{-1} B (Left) or alpha B op(A)^{-1}
 * (Right) with A triangular (uplo, diag).  PlasmaConjTrans is PlasmaTrans.
 * Return codes as the reference (gemm: -1 transa, -2 transb, -3 m, -4 n,
 * -5 k, -8 lda, -10 ldb, -13 ldc; syrk: -1 uplo, -2 trans, -3 n, -4 k,
 * -7 lda, -10 ldc; trsm: -1 side, -2 uplo, -3 transa, -4 diag, -5 m, -6 n,
 * -8 lda, -10 ldb).  Tile forms: descriptors of one square tile size.
 */
int plasma_sgemm(plasma_enum_t transa, plasma_enum_t transb, int m, int n,
                 int k, float alpha, float *pA, int lda, float *pB, int ldb,
                 float beta, float *pC, int ldc);
void plasma_omp_sgemm(plasma_enum_t transa, plasma_enum_t transb,
                      float alpha, plasma_desc_t A, plasma_desc_t B,
                      float beta, plasma_desc_t C, plasma_sequence_t *sequence,
                      plasma_request_t *request);
int plasma_ssyrk(plasma_enum_t uplo, plasma_enum_t trans, int n, int k,
                 float alpha, float *pA, int lda, float beta, float *pC, int ldc);
void plasma_omp_ssyrk(plasma_enum_t uplo, plasma_enum_t trans, float alpha,
                      plasma_desc_t A, float beta, plasma_desc_t C,
                      plasma_sequence_t *sequence, plasma_request_t *request);
int plasma_strsm(plasma_enum_t side, plasma_enum_t uplo,
                 plasma_enum_t transa, plasma_enum_t diag, int m, int n,
                 float alpha, float *pA, int lda, float *pB, int ldb);
void plasma_omp_strsm(plasma_enum_t side, plasma_enum_t uplo,
                      plasma_enum_t transa, plasma_enum_t diag, float alpha,
                      plasma_desc_t A, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request);

/*
 * Band Cholesky (reference compute/zpbtrf.c:77, :200; compute/zpbtrs.c:81,
 * :236; compute/zpbsv.c:92; compute/zpb2desc.c:27; compute/zdesc2pb.c:27).
 * AB: the kd + 1 diagonals of the stored triangle in LAPACK band layout
 * (ldab >= kd + 1), host or device memory; on exit of pbtrf / pbsv the
 * factor in the same layout.  Return codes as the reference (pbtrf: -1
 * uplo, -2 n, -3 kd, -5 ldab; pbtrs: -1, -2, -4 kd, -5 nrhs, -7 ldab,
 * -10 ldb; pbsv: -1, -2, -3 kd, -4 nrhs, -6 ldab, -8 ldb; > 0 the order of
 * the first leading minor that is not positive definite).  Tile forms: a
 * band descriptor (plasma_desc_general_band_create) of the uplo triangle.
 */
int plasma_spbtrf(plasma_enum_t uplo, int n, int kd, float *pAB, int ldab);
int plasma_spbtrs(plasma_enum_t uplo, int n, int kd, int nrhs, float *pAB,
                  int ldab, float *pB, int ldb);
int plasma_spbsv(plasma_enum_t uplo, int n, int kd, int nrhs, float *pAB,
                 int ldab, float *pB, int ldb);
void plasma_omp_spbtrf(plasma_enum_t uplo, plasma_desc_t AB,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_spbtrs(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_spbsv(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_spb2desc(float *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request);
void plasma_omp_sdesc2pb(plasma_desc_t A, float *pA, int lda,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_S_H */
