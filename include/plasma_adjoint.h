/*
 * plasma_adjoint.h -- reverse-mode derivatives of the Cholesky solve path.
 * No counterpart in the reference's interface (include/plasma_z.h has no
 * derivatives); for callers that differentiate through a factorization,
 * such as Gaussian-process log-likelihoods.
 *
 * Gradients are symmetric and written in full: Abar is n x n with
 * df = sum_ij Abar(i,j) dA(i,j) for every symmetric perturbation dA.
 * Arrays are column major and may be host or device memory; the calls
 * return when the result is in place.
 *
 * plasma_?potrf_adjoint: pA holds the Cholesky factor from plasma_?potrf
 * (the uplo triangle; for PlasmaUpper the factor U = L^T) and is not
 * modified.  On entry the uplo triangle of pAbar holds the adjoint of the
 * factor (Ubar = Lbar^T for PlasmaUpper; the other triangle is not read);
 * on exit pAbar holds the symmetric gradient of f(A) = f(L(A)).
 * Returns PlasmaSuccess or -i for an illegal argument i.
 *
 * plasma_?potrs_adjoint: X = A^{-1} B as computed by plasma_?potrs from the
 * factor in pA.  On entry pXbar holds the adjoint of X; on exit it holds
 * the adjoint of B, A^{-1} Xbar, and pAbar (n x n, full) is updated by
 * Abar -= 1/2 (Bbar X^T + X Bbar^T), so that gradients of several terms
 * accumulate.  Returns PlasmaSuccess or -i for an illegal argument i.
 */
#ifndef PLASMA_ADJOINT_H
#define PLASMA_ADJOINT_H

#include "plasma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int plasma_dpotrf_adjoint(plasma_enum_t uplo, int n, const double *pA,
                          int lda, double *pAbar, int ldabar);
int plasma_spotrf_adjoint(plasma_enum_t uplo, int n, const float *pA,
                          int lda, float *pAbar, int ldabar);
int plasma_dpotrs_adjoint(plasma_enum_t uplo, int n, int nrhs,
                          const double *pA, int lda, const double *pX,
                          int ldx, double *pXbar, int ldxbar, double *pAbar,
                          int ldabar);
int plasma_spotrs_adjoint(plasma_enum_t uplo, int n, int nrhs,
                          const float *pA, int lda, const float *pX, int ldx,
                          float *pXbar, int ldxbar, float *pAbar, int ldabar);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_ADJOINT_H */
