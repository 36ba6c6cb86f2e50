"""Pins for the oracle's Cholesky adjoint (no reference counterpart; the
oracle is the reverse-mode sweep of the scalar Cholesky).

* central finite differences (h = 1e-6) of f(A) = <C, L(A)> through the
  oracle's tile potrf, entry by entry, on the lower triangle the
  factorization reads;
* the closed form of the log-determinant: Lbar = diag(2 / L_ii) gives
  Abar = A^{-1};
* the Gaussian-process log-likelihood gradient 1/2 (alpha alpha^T - K^{-1})
  through the Cholesky and potrs adjoints;
* n = 1 and uplo symmetry.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as ora
from tests import _util as U


def _f(A, C, nb=8):
    L, info = ora.potrf(ora.LOWER, A, nb)
    assert info == 0
    return float(np.sum(np.tril(L) * C))


def test_adjoint_finite_differences():
    n, h = 24, 1e-6
    ora.set_blas(ora.BUILTIN)
    A = U.ggt_spd(n, seed=3)
    C = np.tril(np.random.default_rng(4).standard_normal((n, n)))
    L, _ = ora.potrf(ora.LOWER, A, 8)
    G, info = ora.potrf_adjoint(ora.LOWER, np.tril(L), C)
    assert info == 0
    assert np.allclose(G, G.T, rtol=0, atol=0)
    scale = np.max(np.abs(G))
    for i in range(n):
        for j in range(i + 1):
            Ap, Am = A.copy(), A.copy()
            Ap[i, j] += h
            Am[i, j] -= h
            fd = (_f(Ap, C) - _f(Am, C)) / (2 * h)
            # the factorization reads the lower triangle: the derivative
            # along A(i,j) alone is the pair's whole gradient
            want = G[i, j] if i == j else 2.0 * G[i, j]
            assert abs(fd - want) <= 1e-6 * scale, (i, j, fd, want)


@pytest.mark.parametrize("uplo", [ora.LOWER, ora.UPPER])
def test_adjoint_logdet_closed_form(uplo):
    n = 96
    A = U.gp_kernel(n, seed=1)
    F, info = ora.potrf(uplo, A, 16)
    assert info == 0
    Lbar = np.diag(2.0 / np.diag(F))
    G, _ = ora.potrf_adjoint(uplo, F, Lbar)
    assert np.linalg.norm(G @ A - np.eye(n)) < 1e-10


def test_adjoint_upper_is_lower_transposed():
    n = 50
    A = U.ggt_spd(n, seed=5)
    C = np.tril(np.random.default_rng(6).standard_normal((n, n)))
    Fl, _ = ora.potrf(ora.LOWER, A, 8)
    Fu, _ = ora.potrf(ora.UPPER, A, 8)
    Gl, _ = ora.potrf_adjoint(ora.LOWER, Fl, C)
    Gu, _ = ora.potrf_adjoint(ora.UPPER, Fu, C.T.copy())
    assert np.allclose(Gl, Gu, rtol=1e-13, atol=1e-13)


def test_adjoint_order_one():
    a, lb = 6.25, 3.0
    G, _ = ora.potrf_adjoint(ora.LOWER, np.array([[np.sqrt(a)]]),
                             np.array([[lb]]))
    assert G[0, 0] == pytest.approx(lb / (2.0 * np.sqrt(a)), rel=1e-15)


def test_gp_loglik_gradient_closed_form():
    """d/dK [-1/2 y^T K^{-1} y - 1/2 log det K] = 1/2 (a a^T - K^{-1}),
    a = K^{-1} y: the log-determinant term through the Cholesky adjoint
    (Lbar = -diag(1 / L_ii)), the quadratic term through the potrs adjoint
    (Bbar = K^{-1} Xbar, Abar = -1/2 (Bbar X^T + X Bbar^T))."""
    n = 128
    K = U.gp_kernel(n, seed=2)
    y = U.rhs(n, 1, seed=3)
    F, _ = ora.potrf(ora.LOWER, K, 32)
    alpha, _ = ora.potrs(ora.LOWER, F, y, 32)
    G, _ = ora.potrf_adjoint(ora.LOWER, F, -np.diag(1.0 / np.diag(F)))
    bbar, _ = ora.potrs(ora.LOWER, F, -0.5 * y, 32)
    G -= 0.5 * (bbar @ alpha.T + alpha @ bbar.T)
    Kinv = sl.inv(K)
    want = 0.5 * (alpha @ alpha.T - Kinv)
    assert np.linalg.norm(G - want) <= 1e-10 * np.linalg.norm(want)
