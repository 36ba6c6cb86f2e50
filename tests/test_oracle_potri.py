"""The oracle's inverse from the Cholesky factor (CPU only): ora.potri /
ora.poinv, the restatement of the reference's plasma_dpotri /
plasma_dpoinv (compute/zpotri.c:65-186, compute/zpoinv.c:67, tile
algorithms compute/pztrtri.c and compute/pzlauum.c).

Pinned by LAPACK's dpotri (SciPy's LAPACK), by A A^{-1} = I, by the closed
form of a diagonal matrix, by the exact inverse of a factor with integer
entries and unit diagonal (whose inverse is integer too), by LAPACK's info
for a zero on the factor's diagonal, and by the other triangle passing
through unchanged.  Tile sizes that do and do not divide n, both triangles.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle as ora
from tests import _util as U


@pytest.fixture(autouse=True)
def builtin():
    ora.set_blas(ora.BUILTIN)
    yield


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("n,nb", [(1, 4), (37, 8), (64, 16), (150, 32)])
def test_potri_vs_lapack(lower, n, nb):
    A = U.ggt_spd(n, seed=n)
    F, info = ora.potrf(ora.LOWER if lower else ora.UPPER, A, nb)
    assert info == 0
    inv, info = ora.potri(ora.LOWER if lower else ora.UPPER, F, nb)
    assert info == 0
    ref, linfo = la.dpotri(F, lower=lower)
    assert linfo == 0
    tri = np.tril if lower else np.triu
    assert np.linalg.norm(tri(inv) - tri(ref)) <= 1e-13 * np.linalg.norm(
        tri(ref))
    # the other triangle passes through
    other = np.triu(F, 1) if lower else np.tril(F, -1)
    got = np.triu(inv, 1) if lower else np.tril(inv, -1)
    assert np.array_equal(got, other)
    full = U.sym_full(inv, lower)
    assert np.abs(A @ full - np.eye(n)).max() <= 1e-12 * n


@pytest.mark.parametrize("lower", [True, False])
def test_potri_integer_factor_exact(lower):
    """L unit lower triangular with integer entries: L^{-1} is integer and
    so is L^{-T} L^{-1}; the tile algorithm must reproduce it exactly."""
    n, nb = 24, 5
    rng = np.random.default_rng(1)
    L = np.tril(rng.integers(-2, 3, (n, n)).astype(np.float64), -1) + np.eye(n)
    Linv = np.linalg.inv(L).round()
    assert np.array_equal(L @ Linv, np.eye(n))
    exact = Linv.T @ Linv
    F = np.asfortranarray(L if lower else L.T)
    inv, info = ora.potri(ora.LOWER if lower else ora.UPPER, F, nb)
    assert info == 0
    tri = np.tril if lower else np.triu
    assert np.array_equal(tri(inv), tri(exact))


def test_potri_diagonal_closed_form():
    d = np.linspace(0.5, 4.0, 20)
    F = np.asfortranarray(np.diag(d))
    inv, info = ora.potri(ora.LOWER, F, 6)
    assert info == 0
    assert np.allclose(np.diag(inv), 1.0 / d ** 2, rtol=1e-15, atol=0)
    assert np.array_equal(np.tril(inv, -1), np.zeros((20, 20)))


@pytest.mark.parametrize("zero", [0, 7, 19])
def test_potri_singular_info(zero):
    """A zero on the factor's diagonal: info = its 1-based index, as
    LAPACK's dpotri (dtrtri's singularity check)."""
    n, nb = 20, 8
    F = np.asfortranarray(np.linalg.cholesky(U.ggt_spd(n, seed=2)))
    F[zero, zero] = 0.0
    _, info = ora.potri(ora.LOWER, F, nb)
    _, linfo = la.dpotri(F, lower=True)
    assert info == linfo == zero + 1


@pytest.mark.parametrize("lower", [True, False])
def test_poinv_vs_inverse(lower):
    n, nb = 90, 16
    A = U.ggt_spd(n, seed=4)
    inv, info = ora.poinv(ora.LOWER if lower else ora.UPPER, A, nb)
    assert info == 0
    ref = np.linalg.inv(A)
    tri = np.tril if lower else np.triu
    assert np.abs(tri(inv) - tri(ref)).max() <= 1e-13 * np.abs(ref).max()


def test_poinv_not_pd():
    n = 40
    A = U.ggt_spd(n, seed=6)
    A[:, 13] = 0.0
    A[13, :] = 0.0
    _, info = ora.poinv(ora.LOWER, A, 8)
    assert info == 14
