/*
 * plasma_oracle.h -- CPU restatement of PLASMA's tile Cholesky solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline of the MI355X implementation; nothing in the product links
 * or calls it.
 *
 * It restates the reference's algorithms on the reference's own data
 * layout: LAPACK arrays are translated into PLASMA tile storage (the
 * triangular / general descriptors of include/plasma_descriptor.h), the
 * tile algorithms insert one OpenMP task per tile operation with the same
 * depend clauses as the reference (compute/pzpotrf.c, pztrsm.c, zcposv.c
 * ...), and the results are translated back.  Tile kernels are either this
 * file's own loops (ORA_BLAS_BUILTIN) or the vendor BLAS/LAPACK the
 * reference calls (ORA_BLAS_MKL, MKL loaded at run time with dlopen).
 *
 * Parity pins (see tests/test_oracle.py): every routine is checked against
 * an independent LAPACK (scipy / OpenBLAS) on the reference tester's
 * criteria, against closed forms and against brute force on small inputs.
 */
#ifndef PLASMA_ORACLE_H
#define PLASMA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_BLAS_BUILTIN = 0, ORA_BLAS_MKL = 1 };
enum { ORA_UPPER = 121, ORA_LOWER = 122 };

/* Select the tile-kernel backend; returns the backend in use (MKL falls
 * back to built-in when it cannot be loaded). */
int ora_set_blas(int backend);
int ora_get_blas(void);
int ora_mkl_refused(void);
/* inverse from the factor / of A (reference compute/zpotri.c, zpoinv.c) */
int ora_dpotri(int uplo, int n, double *A, int lda, int nb);
int ora_dpoinv(int uplo, int n, double *A, int lda, int nb);
int ora_spotri(int uplo, int n, float *A, int lda, int nb);
/* band Cholesky with tile size nb (reference compute/zpbtrf.c, zpbtrs.c,
 * zpbsv.c): AB in LAPACK band layout (ldab >= kd + 1); ora_xpbsv with
 * factor = 1 factors and solves, with factor = 0 solves with the factor in
 * AB (pbtrs) */
int ora_dpbtrf(int uplo, int n, int kd, double *AB, int ldab, int nb);
int ora_spbtrf(int uplo, int n, int kd, float *AB, int ldab, int nb);
int ora_dpbsv(int uplo, int n, int kd, int nrhs, double *AB, int ldab,
              double *B, int ldb, int nb, int factor);
int ora_spbsv(int uplo, int n, int kd, int nrhs, float *AB, int ldab,
              float *B, int ldb, int nb, int factor);
/* norms (reference compute/zlansy.c, zlange.c, dzamax.c) */
double ora_dlansy(int norm, int uplo, int n, const double *A, int lda);
double ora_dlange(int norm, int m, int n, const double *A, int lda);
void ora_damax(int colrow, int m, int n, const double *A, int lda,
               double *values);
/* Threads used by the OpenMP task runtime (0 = keep the default). */
void ora_set_threads(int nthreads);
int ora_get_threads(void);

/* Test hook: plant a known mistake in the tile Cholesky (1 = drop the
 * syrk update of diagonal tiles).  0 restores the correct algorithm. */
void ora_plant_mistake(int which);

/* plasma_dpotrf / dpotrs / dposv with tile size nb (reference
 * compute/zpotrf.c:74-150, zpotrs.c:77-172, zposv.c:90-186). */
int ora_dpotrf(int uplo, int n, double *A, int lda, int nb);
int ora_dpotrs(int uplo, int n, int nrhs, const double *A, int lda,
               double *B, int ldb, int nb);
int ora_dposv(int uplo, int n, int nrhs, double *A, int lda, double *B,
              int ldb, int nb);
int ora_spotrf(int uplo, int n, float *A, int lda, int nb);
int ora_sposv(int uplo, int n, int nrhs, float *A, int lda, float *B,
              int ldb, int nb);

/* plasma_dsposv (reference compute/zcposv.c:116-271, :363-539). */
int ora_dsposv(int uplo, int n, int nrhs, const double *A, int lda,
               const double *B, int ldb, double *X, int ldx, int *iter,
               int nb);

/* Cholesky adjoint by the scalar reverse-mode sweep (no reference
 * counterpart): the factor in the uplo triangle of L, Lbar in the uplo
 * triangle of Lbar (Ubar = Lbar^T for upper); on exit Lbar holds the full
 * symmetric gradient of f(A) = f(L(A)).  Plain loops, small n only. */
int ora_dpotrf_adjoint(int uplo, int n, const double *L, int ldl,
                       double *Lbar, int ldlb);

/* plasma_dgemm / dsyrk / dtrsm (+ s) with tile size nb (reference
 * compute/zgemm.c:105, zsyrk.c:89, ztrsm.c:104); enums as the reference
 * (111 NoTrans, 112 Trans; 141 Left, 142 Right; 131 NonUnit, 132 Unit). */
int ora_dgemm(int ta, int tb, int m, int n, int k, double alpha,
              const double *A, int lda, const double *B, int ldb, double beta,
              double *C, int ldc, int nb);
int ora_dsyrk(int uplo, int trans, int n, int k, double alpha,
              const double *A, int lda, double beta, double *C, int ldc,
              int nb);
int ora_dtrsm(int side, int uplo, int trans, int diag, int m, int n,
              double alpha, const double *A, int lda, double *B, int ldb,
              int nb);
int ora_sgemm(int ta, int tb, int m, int n, int k, float alpha,
              const float *A, int lda, const float *B, int ldb, float beta,
              float *C, int ldc, int nb);
int ora_strsm(int side, int uplo, int trans, int diag, int m, int n,
              float alpha, const float *A, int lda, float *B, int ldb, int nb);

/* plasma_dlag2s / plasma_slag2d on LAPACK arrays. */
int ora_dlag2s(int m, int n, const double *A, int lda, float *As, int ldas);
void ora_slag2d(int m, int n, const float *As, int ldas, double *A, int lda);

/* Tile layout round trip: copy the uplo (or general when uplo is neither)
 * tiles of A into PLASMA tile storage and write the storage image to
 * `tiles` (size from ora_tile_storage_elems), then translate back into B. */
long ora_tile_storage_elems(int uplo, int m, int n, int nb);
void ora_lapack2tile(int uplo, int m, int n, const double *A, int lda,
                     double *tiles, int nb);
void ora_tile2lapack(int uplo, int m, int n, const double *tiles, double *A,
                     int lda, int nb);

#ifdef __cplusplus
}
#endif

/* Complex precisions (void * = interleaved double complex for z, float
 * complex for c): Hermitian Cholesky, solves and the mixed solver. */
int ora_zpotrf(int uplo, int n, void *A, int lda, int nb);
int ora_cpotrf(int uplo, int n, void *A, int lda, int nb);
int ora_zposv(int uplo, int n, int nrhs, void *A, int lda, void *B, int ldb,
              int nb);
int ora_cposv(int uplo, int n, int nrhs, void *A, int lda, void *B, int ldb,
              int nb);
int ora_zpotrs(int uplo, int n, int nrhs, const void *L, int lda, void *B,
               int ldb, int nb);
int ora_cpotrs(int uplo, int n, int nrhs, const void *L, int lda, void *B,
               int ldb, int nb);
int ora_zcposv(int uplo, int n, int nrhs, const void *A, int lda,
               const void *B, int ldb, void *X, int ldx, int *iter, int nb);
int ora_zlag2c(int m, int n, const void *A, int lda, void *As, int ldas);
void ora_clag2z(int m, int n, const void *As, int ldas, void *A, int lda);
double ora_zlanhe_inf(int uplo, int n, const void *A, int lda);
void ora_dzamax(int m, int n, const void *A, int lda, double *values);

#endif
