"""The oracle's band Cholesky (CPU only): ora.pbtrf / ora.pbsv, the
restatement of the reference's plasma_dpbtrf / plasma_dpbtrs / plasma_dpbsv
(compute/zpbtrf.c, zpbtrs.c, zpbsv.c; tile algorithms compute/pzpbtrf.c
and the Left variants of compute/pztbsm.c).

Pinned by LAPACK's dpbtrf / dpbtrs (SciPy), by the exact factor of a band
matrix built from an integer band factor (every intermediate value an
exactly represented integer), by the diagonal closed form (kd = 0), by
LAPACK's info for a leading minor that is not positive definite, and by the
band array's rows outside the band passing through.  Band widths below,
at and above the tile size, n not a multiple of nb, both triangles, both
tile-kernel backends."""
import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle as ora
from tests import _band as BD
from tests import _util as U


@pytest.fixture(params=["builtin", "mkl"])
def backend(request):
    if request.param == "mkl":
        if ora.set_blas(ora.MKL) != ora.MKL:
            pytest.skip("MKL not available")
    else:
        ora.set_blas(ora.BUILTIN)
    yield request.param
    ora.set_blas(ora.BUILTIN)


CASES = [(1, 0, 4), (50, 0, 8), (61, 1, 8), (100, 7, 16), (130, 16, 16),
         (131, 17, 16), (200, 45, 16), (97, 96, 32)]


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("n,kd,nb", CASES)
def test_pbtrf_vs_lapack(backend, lower, n, kd, nb):
    A = BD.spd_band(n, kd, seed=n + kd)
    ab = BD.to_band(A, kd, lower)
    f, info = ora.pbtrf(ora.LOWER if lower else ora.UPPER, ab, kd, nb)
    ref, linfo = la.dpbtrf(ab, lower=int(lower))
    assert info == linfo == 0
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max()
    F = BD.from_band(f, kd, lower)
    R = F @ F.T if lower else F.T @ F
    assert np.abs(R - A).max() <= 1e-13 * np.abs(A).max()


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("n,kd,nb", [(70, 5, 8), (130, 20, 16), (64, 63, 16)])
def test_pbsv_pbtrs_vs_lapack(backend, lower, n, kd, nb):
    A = BD.spd_band(n, kd, seed=3)
    ab = BD.to_band(A, kd, lower)
    B = U.rhs(n, 5, seed=4)
    f, X, info = ora.pbsv(ora.LOWER if lower else ora.UPPER, ab, kd, B, nb)
    assert info == 0
    ref, linfo = la.dpbtrf(ab, lower=int(lower))
    Xref, linfo2 = la.dpbtrs(ref, B, lower=int(lower))
    assert linfo == linfo2 == 0
    assert np.abs(X - Xref).max() <= 1e-12 * np.abs(Xref).max()
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    # pbtrs from the factor gives the same solution
    _, X2, info = ora.pbsv(ora.LOWER if lower else ora.UPPER, f, kd, B, nb,
                           factor=False)
    assert info == 0 and np.array_equal(X2, X)


@pytest.mark.parametrize("lower", [True, False])
def test_exact_integer_factor(backend, lower):
    """A = L L^T from an integer band factor with positive integer
    diagonal: every operation of the factorization is exact, so the factor
    comes back bit for bit."""
    n, kd, nb = 90, 13, 16
    rng = np.random.default_rng(7)
    L = np.zeros((n, n))
    for d in range(1, kd + 1):
        L[np.arange(d, n), np.arange(n - d)] = rng.integers(-3, 4, n - d)
    L[np.diag_indices(n)] = rng.integers(1, 4, n)
    A = L @ L.T
    ab = BD.to_band(A, kd, lower)
    f, info = ora.pbtrf(ora.LOWER if lower else ora.UPPER, ab, kd, nb)
    assert info == 0
    want = BD.to_band(L if lower else L.T, kd, lower)
    assert np.array_equal(f, want)


def test_diagonal_closed_form(backend):
    n = 40
    d = np.arange(1, n + 1, dtype=float) ** 2
    ab = d.reshape(1, n).copy(order="F")
    f, info = ora.pbtrf(ora.LOWER, ab, 0, 8)
    assert info == 0 and np.array_equal(f[0], np.arange(1, n + 1.0))


@pytest.mark.parametrize("lower", [True, False])
def test_not_pd_info(backend, lower):
    n, kd, nb = 120, 9, 16
    A = BD.spd_band(n, kd, seed=1)
    A[57, 57] = -50.0
    ab = BD.to_band(A, kd, lower)
    f, info = ora.pbtrf(ora.LOWER if lower else ora.UPPER, ab, kd, nb)
    _, linfo = la.dpbtrf(ab, lower=int(lower))
    assert info == linfo > 0
    assert np.array_equal(f, ab)          # not written back on failure


def test_rows_outside_band_pass_through(backend):
    n, kd, nb = 50, 4, 8
    A = BD.spd_band(n, kd, seed=2)
    ab = BD.to_band(A, kd, True, ldab=kd + 3)
    ab[kd + 1:, :] = 99.0                  # padding rows of the array
    f, info = ora.pbtrf(ora.LOWER, ab, kd, nb)
    assert info == 0 and np.all(f[kd + 1:, :] == 99.0)
