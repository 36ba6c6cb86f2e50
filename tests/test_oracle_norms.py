"""The oracle's norms (CPU only): ora.lansy / ora.lange / ora.amax, the
restatement of the reference's plasma_dlansy / plasma_dlange / plasma_damax
(compute/zlansy.c:81, compute/zlange.c:76, compute/dzamax.c).

Pinned by LAPACK's own dlange (SciPy's OpenBLAS) and dlansy (MKL's
LAPACKE, when present), by NumPy closed forms on the full matrix, and by special
cases that a plausible mistake breaks: the stored triangle of a symmetric
matrix only (the other triangle holds garbage), Frobenius without overflow
at entries near 1e200, NaN propagation, a rectangular matrix (One and Inf
swap roles), padding (lda > m) and empty matrices."""
import ctypes

import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle as ora
from tests import _util as U

NORMS = {"M": ora.MAXNORM, "1": ora.ONENORM, "I": ora.INFNORM,
         "F": ora.FRONORM}


def _lapacke_dlansy(norm, uplo, a):
    """MKL's LAPACKE_dlansy (column major), or None without MKL."""
    m = U.mkl()
    if m is None:
        return None
    f = m.LAPACKE_dlansy
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_int, ctypes.c_char, ctypes.c_char, ctypes.c_int,
                  ctypes.c_void_p, ctypes.c_int]
    return f(102, norm.encode(), uplo.encode(), a.shape[0], a.ctypes.data,
             max(1, a.shape[0]))


def _np_norm(norm, a):
    if norm == "M":
        return np.abs(a).max()
    if norm == "1":
        return np.abs(a).sum(axis=0).max()
    if norm == "I":
        return np.abs(a).sum(axis=1).max()
    return np.sqrt((a * a).sum())


@pytest.mark.parametrize("norm", list(NORMS))
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("n", [1, 7, 130])
def test_lansy_vs_lapack(norm, lower, n):
    rng = np.random.default_rng(n)
    S = rng.standard_normal((n, n))
    S = S + S.T
    stored = np.asfortranarray(np.tril(S) if lower else np.triu(S))
    # the other triangle holds garbage that must not be read
    junk = 1e30 * (np.triu(np.ones((n, n)), 1) if lower
                   else np.tril(np.ones((n, n)), -1))
    A = np.asfortranarray(stored + junk)
    v = ora.lansy(NORMS[norm], ora.LOWER if lower else ora.UPPER, A)
    ref = _np_norm(norm, S)
    assert abs(v - ref) <= 1e-13 * ref
    lap = _lapacke_dlansy(norm, "L" if lower else "U", A)
    if lap is not None:
        assert abs(v - lap) <= 1e-14 * lap


@pytest.mark.parametrize("norm", list(NORMS))
@pytest.mark.parametrize("m,n", [(1, 1), (9, 4), (4, 9), (200, 131)])
def test_lange_vs_lapack(norm, m, n):
    rng = np.random.default_rng(m * n)
    P = np.asfortranarray(rng.standard_normal((m + 3, n)))   # lda = m + 3
    A = P[:m]
    v = ora.lange(NORMS[norm], A)
    ref = la.dlange(norm, np.asfortranarray(A))
    assert abs(v - ref) <= 1e-14 * ref
    assert abs(v - _np_norm(norm, A)) <= 1e-13 * ref


def test_frobenius_no_overflow_and_nan():
    A = np.asfortranarray(np.full((5, 4), 1e200))
    assert np.isclose(ora.lange(ora.FRONORM, A), 1e200 * np.sqrt(20.0),
                      rtol=1e-14)
    B = np.asfortranarray(np.ones((3, 3)))
    B[1, 2] = np.nan
    for norm in NORMS.values():
        assert np.isnan(ora.lange(norm, B))


def test_empty():
    z = np.zeros((0, 0), order="F")
    for norm in NORMS.values():
        assert ora.lange(norm, z) == 0.0
        assert ora.lansy(norm, ora.LOWER, z) == 0.0


@pytest.mark.parametrize("m,n", [(1, 1), (17, 5), (300, 64)])
def test_amax(m, n):
    rng = np.random.default_rng(m + n)
    P = np.asfortranarray(rng.standard_normal((m + 2, n)))
    A = P[:m]
    assert np.array_equal(ora.amax(ora.COLUMNWISE, A), np.abs(A).max(axis=0))
    assert np.array_equal(ora.amax(ora.ROWWISE, A), np.abs(A).max(axis=1))
