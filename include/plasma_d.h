/*
 * plasma_d.h -- double precision Cholesky solve path.
 *
 * Signatures are the reference's d-precision forms (generated from
 * include/plasma_z.h:272-292, :398-430, :631-646, :667-671).
 *
 * LAPACK-layout routines (plasma_dpotrf, plasma_dpotrs, plasma_dposv) take
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
#ifndef PLASMA_D_H
#define PLASMA_D_H

#include "plasma_async.h"
#include "plasma_descriptor.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference compute/zpotrf.c:74 */
int plasma_dpotrf(plasma_enum_t uplo, int n, double *pA, int lda);
/* reference compute/zpotrs.c:77 */
int plasma_dpotrs(plasma_enum_t uplo, int n, int nrhs,
                  double *pA, int lda, double *pB, int ldb);
/* reference compute/zposv.c:90 */
int plasma_dposv(plasma_enum_t uplo, int n, int nrhs,
                 double *pA, int lda, double *pB, int ldb);

/* reference compute/zpotrf.c:206 */
void plasma_omp_dpotrf(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);
/* reference compute/zpotrs.c:226 */
void plasma_omp_dpotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request);
/* reference compute/zposv.c:244 */
void plasma_omp_dposv(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request);

/* reference compute/zge2desc.c:27, zdesc2ge.c, ztr2desc.c, zdesc2tr.c */
void plasma_omp_dge2desc(double *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_ddesc2ge(plasma_desc_t A, double *pA, int lda,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_dtr2desc(double *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_ddesc2tr(plasma_desc_t A, double *pA, int lda,
                         plasma_sequence_t *sequence, plasma_request_t *request);

/* Inverse of an SPD matrix: from its Cholesky factor (reference
 * compute/zpotri.c:65, :206) or from the matrix itself (compute/zpoinv.c:67,
 * :201); pA / A end with A^{-1} in the uplo triangle.  Return codes as
 * plasma_dpotrf (-1 uplo, -2 n, -4 lda; i > 0: a zero at position i of the
 * factor's diagonal, or potrf's info). */
int plasma_dpotri(plasma_enum_t uplo, int n, double *pA, int lda);
int plasma_dpoinv(plasma_enum_t uplo, int n, double *pA, int lda);
void plasma_omp_dpotri(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_dpoinv(plasma_enum_t uplo, plasma_desc_t A,
                       plasma_sequence_t *sequence, plasma_request_t *request);

/*
 * Norms (reference compute/zlansy.c:81, :240; compute/zlange.c:76, :233;
 * compute/dzamax.c:21, :113).  norm: PlasmaMaxNorm, PlasmaOneNorm,
 * PlasmaInfNorm or PlasmaFrobeniusNorm; colrow: PlasmaColumnwise (values
 * has n entries) or PlasmaRowwise (m entries).  The synchronous forms
 * return the norm, or a negative argument code.  Tile forms: work is device
 * memory of plasma_norm_work_doubles(A) doubles or NULL (library
 * workspace); value / values may be host or device memory and hold the
 * result after plasma_sequence_sync; views start on a tile boundary.
 */
double plasma_dlansy(plasma_enum_t norm, plasma_enum_t uplo, int n,
                     double *pA, int lda);
double plasma_dlange(plasma_enum_t norm, int m, int n, double *pA, int lda);
int plasma_damax(plasma_enum_t colrow, int m, int n, double *pA, int lda,
                 double *values);
void plasma_omp_dlansy(plasma_enum_t norm, plasma_enum_t uplo,
                       plasma_desc_t A, double *work, double *value,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_dlange(plasma_enum_t norm, plasma_desc_t A, double *work,
                       double *value, plasma_sequence_t *sequence,
                       plasma_request_t *request);
void plasma_omp_damax(plasma_enum_t colrow, plasma_desc_t A, double *work,
                      double *values, plasma_sequence_t *sequence,
                      plasma_request_t *request);
/* No reference counterpart: device workspace (doubles) of the tile norms. */
size_t plasma_norm_work_doubles(plasma_desc_t A);

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
int plasma_dgemm(plasma_enum_t transa, plasma_enum_t transb, int m, int n,
                 int k, double alpha, double *pA, int lda, double *pB, int ldb,
                 double beta, double *pC, int ldc);
void plasma_omp_dgemm(plasma_enum_t transa, plasma_enum_t transb,
                      double alpha, plasma_desc_t A, plasma_desc_t B,
                      double beta, plasma_desc_t C, plasma_sequence_t *sequence,
                      plasma_request_t *request);
int plasma_dsyrk(plasma_enum_t uplo, plasma_enum_t trans, int n, int k,
                 double alpha, double *pA, int lda, double beta, double *pC, int ldc);
void plasma_omp_dsyrk(plasma_enum_t uplo, plasma_enum_t trans, double alpha,
                      plasma_desc_t A, double beta, plasma_desc_t C,
                      plasma_sequence_t *sequence, plasma_request_t *request);
int plasma_dtrsm(plasma_enum_t side, plasma_enum_t uplo,
                 plasma_enum_t transa, plasma_enum_t diag, int m, int n,
                 double alpha, double *pA, int lda, double *pB, int ldb);
void plasma_omp_dtrsm(plasma_enum_t side, plasma_enum_t uplo,
                      plasma_enum_t transa, plasma_enum_t diag, double alpha,
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
int plasma_dpbtrf(plasma_enum_t uplo, int n, int kd, double *pAB, int ldab);
int plasma_dpbtrs(plasma_enum_t uplo, int n, int kd, int nrhs, double *pAB,
                  int ldab, double *pB, int ldb);
int plasma_dpbsv(plasma_enum_t uplo, int n, int kd, int nrhs, double *pAB,
                 int ldab, double *pB, int ldb);
void plasma_omp_dpbtrf(plasma_enum_t uplo, plasma_desc_t AB,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_dpbtrs(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_dpbsv(plasma_enum_t uplo, plasma_desc_t AB, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_dpb2desc(double *pA, int lda, plasma_desc_t A,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request);
void plasma_omp_ddesc2pb(plasma_desc_t A, double *pA, int lda,
                         plasma_sequence_t *sequence,
                         plasma_request_t *request);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_D_H */
