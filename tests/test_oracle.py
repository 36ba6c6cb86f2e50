"""CPU tests that pin the oracle (oracle/liboracle.so) before it is trusted.

The reference ships no golden vectors (its tester compares against LAPACK at
run time, test/test_zpotrf.c:124-152, test_zposv.c:144-168,
test_zcposv.c:327-373), so the oracle is pinned the same way plus by what
the mathematics fixes:
  * an independent LAPACK (scipy's OpenBLAS dpotrf / dposv) on the tester's
    criteria, for both oracle tile-kernel backends (built-in loops, MKL);
  * MKL's LAPACKE_dsposv for the mixed-precision iteration count and the
    solution (same algorithm and stopping rule as zcposv.c);
  * closed forms: diagonal matrices, exact integer factors A = L0 L0^T
    (Cholesky reproduces L0 bit for bit), 2x2 matrices;
  * brute force in extended precision for n <= 16;
  * IEEE round-to-nearest for the FP64 <-> FP32 conversion, and the
    dlag2s overflow rule;
  * the tile layout against an independent NumPy slicing of the matrix;
  * a planted mistake (the syrk of diagonal tiles dropped) must fail.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as ora
from tests import _util as U

BACKENDS = [ora.BUILTIN, ora.MKL]


@pytest.fixture(params=BACKENDS, ids=["builtin", "mkl"])
def backend(request):
    got = ora.set_blas(request.param)
    if got != request.param:
        pytest.skip("MKL not loadable on this host")
    yield got
    ora.set_blas(ora.BUILTIN)


@pytest.mark.parametrize("uplo", [ora.LOWER, ora.UPPER])
@pytest.mark.parametrize("n,nb", [(1, 4), (7, 4), (100, 16), (257, 64),
                                  (300, 128), (512, 128)])
def test_potrf_vs_lapack(backend, uplo, n, nb):
    A = U.tester_spd(n)
    F, info = ora.potrf(uplo, A, nb)
    R, rinfo = sl.lapack.dpotrf(A, lower=(uplo == ora.LOWER), clean=False)
    assert info == rinfo == 0
    assert U.factor_error(F, R, uplo == ora.LOWER) < 50 * U.EPS64
    other = np.triu(F, 1) if uplo == ora.LOWER else np.tril(F, -1)
    same = np.triu(A, 1) if uplo == ora.LOWER else np.tril(A, -1)
    assert np.array_equal(other, same)  # untouched triangle


@pytest.mark.parametrize("uplo", [ora.LOWER, ora.UPPER])
@pytest.mark.parametrize("zerocol", [0, 5, 63, 64, 199])
def test_potrf_info_and_untouched_on_failure(backend, uplo, zerocol):
    n, nb = 200, 64
    A = U.tester_spd(n)
    A[:, zerocol] = 0.0
    F, info = ora.potrf(uplo, A, nb)
    _, rinfo = sl.lapack.dpotrf(A, lower=(uplo == ora.LOWER))
    assert info == rinfo == zerocol + 1
    assert np.array_equal(F, A)  # the failed sequence skips desc2tr


@pytest.mark.parametrize("uplo", [ora.LOWER, ora.UPPER])
@pytest.mark.parametrize("n,nb,nrhs", [(50, 16, 3), (333, 64, 17),
                                       (512, 256, 64)])
def test_posv_potrs(backend, uplo, n, nb, nrhs):
    A = U.ggt_spd(n, seed=n)
    B = U.rhs(n, nrhs)
    F, X, info = ora.posv(uplo, A, B, nb)
    assert info == 0
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    Xs = sl.solve(A, B, assume_a="pos")
    assert np.abs(X - Xs).max() <= 1e-12 * np.abs(Xs).max()
    X2, info2 = ora.potrs(uplo, F, B, nb)
    assert info2 == 0
    assert np.abs(X2 - X).max() <= 1e-13 * np.abs(X).max()


def test_closed_forms(backend):
    d = np.array([4.0, 9.0, 2.25, 1e-4, 1e6])
    F, info = ora.potrf(ora.LOWER, np.diag(d), 2)
    assert info == 0 and np.array_equal(np.diag(F), np.sqrt(d))
    # 2x2: [[a, b], [b, c]] -> l11 = sqrt a, l21 = b / l11, l22 = sqrt(c - l21^2)
    a, b, c = 4.0, 2.0, 5.0
    F, _ = ora.potrf(ora.LOWER, np.array([[a, b], [b, c]]), 1)
    assert np.array_equal(np.tril(F), np.array([[2.0, 0.0], [1.0, 2.0]]))


@pytest.mark.parametrize("uplo", [ora.LOWER, ora.UPPER])
@pytest.mark.parametrize("n,nb", [(6, 2), (24, 5), (40, 16)])
def test_exact_integer_factor(backend, uplo, n, nb):
    """A = L0 L0^T with small integer L0: every operation is exact, so the
    computed factor must equal L0 bit for bit (a dropped term, a wrong
    index or a transposed operand changes it)."""
    rng = np.random.default_rng(n)
    L0 = np.tril(rng.integers(-3, 4, size=(n, n))).astype(np.float64)
    # power-of-two pivots: exact whether a backend divides or multiplies
    # by the reciprocal
    np.fill_diagonal(L0, 2.0 ** rng.integers(0, 3, size=n))
    A = L0 @ L0.T
    assert np.abs(A).max() < 2 ** 40
    F, info = ora.potrf(uplo, A, nb)
    assert info == 0
    if uplo == ora.LOWER:
        assert np.array_equal(np.tril(F), L0)
    else:
        assert np.array_equal(np.triu(F), L0.T)


@pytest.mark.parametrize("n", [3, 9, 16])
def test_brute_force_extended_precision(backend, n):
    A = U.tester_spd(n).astype(np.longdouble)
    L = np.zeros_like(A)
    for j in range(n):
        s = A[j, j] - np.sum(L[j, :j] ** 2)
        L[j, j] = np.sqrt(s)
        for i in range(j + 1, n):
            L[i, j] = (A[i, j] - np.sum(L[i, :j] * L[j, :j])) / L[j, j]
    F, _ = ora.potrf(ora.LOWER, A.astype(np.float64), 4)
    err = np.abs(np.tril(F) - L.astype(np.float64)).max() / np.abs(L).max()
    assert err < 4 * U.EPS64


def test_spotrf_vs_lapack(backend):
    n, nb = 300, 64
    A = U.ggt_spd(n, dtype=np.float32)
    F, info = ora.potrf(ora.LOWER, A, nb, dtype=np.float32)
    R, rinfo = sl.lapack.spotrf(A, lower=True, clean=False)
    assert info == rinfo == 0
    assert U.factor_error(F, R, True) < 50 * U.EPS32


def _mkl_dsposv(A, B, lower=True):
    m = U.mkl()
    if m is None:
        pytest.skip("MKL not available for LAPACKE_dsposv")
    n, nrhs = B.shape
    a = np.asfortranarray(A.copy())
    b = np.asfortranarray(B.copy())
    x = np.zeros_like(b, order="F")
    it = ctypes.c_int()
    vp = ctypes.c_void_p
    info = m.LAPACKE_dsposv(102, ord("L" if lower else "U"), n, nrhs,
                            vp(a.ctypes.data), n, vp(b.ctypes.data), n,
                            vp(x.ctypes.data), n, ctypes.byref(it))
    return x, it.value, info


def _cond_spd(n, kappa, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = kappa ** (-np.arange(n) / (n - 1.0))
    return np.asfortranarray((Q * s) @ Q.T)


@pytest.mark.parametrize("kappa", [1e1, 1e3, 1e5])
@pytest.mark.parametrize("uplo", [ora.LOWER, ora.UPPER])
def test_dsposv_vs_lapack(backend, kappa, uplo):
    n, nb, nrhs = 256, 64, 4
    A = _cond_spd(n, kappa, seed=int(kappa))
    B = U.rhs(n, nrhs)
    X, it, info = ora.dsposv(uplo, A, B, nb)
    assert info == 0 and 0 <= it <= 30
    Xm, itm, infom = _mkl_dsposv(A, B, lower=(uplo == ora.LOWER))
    assert infom == 0
    # same algorithm and stopping rule: iteration counts agree within one
    # (the FP32 factors differ in rounding between tiles and LAPACK)
    assert abs(it - itm) <= 1
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64 * np.sqrt(n)
    assert np.abs(X - Xm).max() <= 1e-12 * kappa * np.abs(Xm).max()


def test_dsposv_fp64_fallback_and_breakdown(backend):
    """kappa beyond FP32: the FP32 factor breaks down; the reference then
    fails the sequence with the FP32 info and does not fall back
    (compute/zcposv.c:445-452 with compute/pzpotrf.c:31-32)."""
    n, nb = 128, 32
    A = _cond_spd(n, 1e12, seed=3)
    B = U.rhs(n, 2)
    X, it, info = ora.dsposv(ora.LOWER, A, B, nb)
    _, sinfo = sl.lapack.spotrf(A.astype(np.float32), lower=True)
    if sinfo > 0:
        assert info > 0
    else:
        assert info == 0 and it in (-31,) + tuple(range(31))


def test_lag2_conversions():
    rng = np.random.default_rng(0)
    A = np.asfortranarray(rng.standard_normal((37, 11)) * 1e3)
    S, info = ora.dlag2s(A)
    assert info == 0
    assert np.array_equal(S, A.astype(np.float32))  # IEEE round to nearest
    assert np.array_equal(ora.slag2d(S), S.astype(np.float64))
    for bad in (1e39, -1e39, np.inf):
        B = A.copy(order="F")
        B[5, 3] = bad
        assert ora.dlag2s(B)[1] == 1
    B = A.copy(order="F")
    B[0, 0] = np.finfo(np.float32).max  # representable: no overflow
    assert ora.dlag2s(B)[1] == 0


def _tile_image(A, nb, region):
    """PLASMA tile storage built by NumPy slicing: full tiles column by
    column, then the short last tile row, the short last tile column and
    the corner tile; triangular storage keeps only the region's tiles."""
    m, n = A.shape
    lm1, ln1 = m // nb, n // nb
    keep = {ora.LOWER: lambda i, j: i >= j, ora.UPPER: lambda i, j: i <= j,
            0: lambda i, j: True}[region]
    parts = []
    for j in range(ln1):
        for i in range(lm1):
            if keep(i, j):
                parts.append(A[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb].ravel("F"))
    if region == 0:
        for j in range(ln1):                                   # A21
            parts.append(A[lm1 * nb:, j * nb:(j + 1) * nb].ravel("F"))
        for i in range(lm1):                                   # A12
            parts.append(A[i * nb:(i + 1) * nb, ln1 * nb:].ravel("F"))
        parts.append(A[lm1 * nb:, ln1 * nb:].ravel("F"))        # A22
    return np.concatenate(parts) if parts else np.zeros(0)


@pytest.mark.parametrize("m,n,nb", [(64, 64, 16), (100, 100, 16),
                                    (70, 45, 16), (257, 257, 64)])
@pytest.mark.parametrize("region", [ora.LOWER, ora.UPPER, 0])
def test_tile_layout(m, n, nb, region):
    if region and m != n:
        pytest.skip("triangular descriptors are square")
    A = np.asfortranarray(np.random.default_rng(m + n).standard_normal((m, n)))
    img = ora.lapack2tile(region, A, nb)
    want = _tile_image(A, nb, region)
    if region == 0 or (m % nb == 0):
        assert np.array_equal(img[:want.size], want)
    else:
        assert np.array_equal(img[:want.size], want)  # full-tile part
    back = ora.tile2lapack(region, img, m, n, nb, init=np.zeros((m, n)))
    mask = np.ones((m, n), bool)
    if region:
        ti = np.arange(m)[:, None] // nb
        tj = np.arange(n)[None, :] // nb
        mask = (ti >= tj) if region == ora.LOWER else (ti <= tj)
    assert np.array_equal(back[mask], A[mask])
    assert np.all(back[~mask] == 0)


def test_planted_mistake_is_caught():
    """Dropping the syrk update of diagonal tiles (a plausible slip) must
    break the factor check; restoring the algorithm fixes it."""
    A = U.tester_spd(200)
    R = sl.lapack.dpotrf(A, lower=True, clean=False)[0]
    ora.set_blas(ora.BUILTIN)
    try:
        ora.plant_mistake(1)
        F, info = ora.potrf(ora.LOWER, A, 32)
        bad = info != 0 or U.factor_error(F, R, True) > 50 * U.EPS64
        assert bad
    finally:
        ora.plant_mistake(0)
    F, info = ora.potrf(ora.LOWER, A, 32)
    assert info == 0 and U.factor_error(F, R, True) < 50 * U.EPS64
