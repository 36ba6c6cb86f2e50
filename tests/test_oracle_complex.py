"""The oracle's complex (Hermitian) restatement pinned against LAPACK
(scipy's zpotrf / cpotrf / zpotrs), numpy and closed forms, on the CPU.

Reference: compute/pzpotrf.c:27-127 and pztrsm.c:43-177 with the z tile
kernels, compute/zcposv.c:363-539, core_zlag2c.c (LAPACK zlag2c), pzlanhe.c
and core_dzamax.c.  A planted mistake (no conjugation in the herk tasks)
must fail the factor check."""
import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle as ora

L, U = ora.LOWER, ora.UPPER


def herm_pd(n, seed=0, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = g @ g.conj().T / n + np.eye(n)
    return a.astype(dtype)


def herm_cond(n, kappa, seed=0):
    """A = Q diag(s) Q^H, Q unitary, s log-spaced on [1/kappa, 1]."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q, _ = np.linalg.qr(g)
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    a = (q * s) @ q.conj().T
    return (a + a.conj().T) / 2


def tri(a, lower):
    return np.tril(a) if lower else np.triu(a)


@pytest.fixture(autouse=True)
def _no_plant():
    ora.plant_mistake(0)
    yield
    ora.plant_mistake(0)


@pytest.mark.parametrize("uplo", [L, U])
@pytest.mark.parametrize("n,nb", [(1, 4), (37, 8), (64, 16), (130, 32),
                                  (203, 64)])
def test_zpotrf_matches_lapack(uplo, n, nb):
    a = herm_pd(n, seed=n)
    f, info = ora.zpotrf(uplo, a, nb)
    assert info == 0
    ref, rinfo = la.zpotrf(a, lower=uplo == L, clean=False)
    assert rinfo == 0
    lo = uplo == L
    err = np.linalg.norm(tri(f, lo) - tri(ref, lo)) / np.linalg.norm(tri(ref, lo))
    assert err < 1e-14, err
    # the factor's diagonal is real; the other triangle is untouched
    assert np.all(np.diag(f).imag == 0)
    other = np.triu(a, 1) if lo else np.tril(a, -1)
    assert np.array_equal(np.triu(f, 1) if lo else np.tril(f, -1), other)


@pytest.mark.parametrize("uplo", [L, U])
def test_cpotrf_matches_lapack(uplo):
    n, nb = 150, 32
    a = herm_pd(n, seed=3, dtype=np.complex64)
    f, info = ora.zpotrf(uplo, a, nb, dtype=np.complex64)
    assert info == 0
    ref, rinfo = la.cpotrf(a, lower=uplo == L, clean=False)
    assert rinfo == 0
    lo = uplo == L
    err = np.linalg.norm(tri(f, lo) - tri(ref, lo)) / np.linalg.norm(tri(ref, lo))
    assert err < 2e-6, err


def test_reconstruction_and_diagonal_imaginary_part_ignored():
    """L L^H = A; a nonzero imaginary part on A's diagonal is not read
    (LAPACK zpotrf uses the real part)."""
    n, nb = 96, 16
    a = herm_pd(n, seed=7)
    a2 = a.copy()
    a2[np.diag_indices(n)] += 1j * np.linspace(1, 2, n)
    f, info = ora.zpotrf(L, a2, nb)
    assert info == 0
    l = np.tril(f)
    assert np.abs(l @ l.conj().T - a).max() < 1e-13


def test_planted_missing_conjugation_is_caught():
    n, nb = 64, 16
    a = herm_pd(n, seed=9)
    ref, _ = la.zpotrf(a, lower=True, clean=True)
    ora.plant_mistake(5)
    f, info = ora.zpotrf(L, a, nb)
    bad = info != 0 or np.linalg.norm(np.tril(f) - ref) / np.linalg.norm(ref) > 1e-6
    assert bad


def test_real_input_equals_real_factor():
    """A real symmetric matrix as complex data: the complex factor is the
    real one (closed form of the embedding)."""
    n, nb = 80, 16
    rng = np.random.default_rng(4)
    g = rng.standard_normal((n, n))
    a = g @ g.T / n + np.eye(n)
    fz, info = ora.zpotrf(L, a.astype(np.complex128), nb)
    fd, infod = ora.potrf(L, a, nb)
    assert info == 0 and infod == 0
    assert np.abs(np.tril(fz).imag).max() == 0
    assert np.abs(np.tril(fz).real - np.tril(fd)).max() < 1e-13


@pytest.mark.parametrize("uplo", [L, U])
def test_not_positive_definite_info(uplo):
    """Leading minor of order j0 + 1 not PD: info = j0 + 1 (1-based,
    global), as LAPACK."""
    n, nb, j0 = 70, 16, 37
    a = herm_pd(n, seed=11)
    a[j0, j0] = -5.0
    _, info = ora.zpotrf(uplo, a, nb)
    _, rinfo = la.zpotrf(a, lower=uplo == L)
    assert info == rinfo == j0 + 1


@pytest.mark.parametrize("uplo", [L, U])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12),
                                       (np.complex64, 1e-4)])
def test_posv_and_potrs(uplo, dtype, tol):
    n, nrhs, nb = 110, 7, 32
    a = herm_pd(n, seed=5, dtype=dtype)
    rng = np.random.default_rng(6)
    b = (rng.standard_normal((n, nrhs))
         + 1j * rng.standard_normal((n, nrhs))).astype(dtype)
    f, x, info = ora.zposv(uplo, a, b, nb, dtype=dtype)
    assert info == 0
    want = np.linalg.solve(a.astype(np.complex128), b.astype(np.complex128))
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < tol
    x2, info = ora.zpotrs(uplo, f, b, nb, dtype=dtype)
    assert info == 0
    assert np.array_equal(x2, x)
    if dtype == np.complex128:
        ref, _ = la.zpotrf(a, lower=uplo == L, clean=False)
        xr, _ = la.zpotrs(ref, b, lower=uplo == L)
        assert np.linalg.norm(x2 - xr) / np.linalg.norm(xr) < 1e-13


@pytest.mark.parametrize("uplo", [L, U])
@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_zcposv(uplo, kappa):
    """FP64 backward error from the single complex factor: the tester's
    scaled residual, X against numpy's solve within cond * eps, and a small
    iteration count."""
    n, nrhs, nb = 160, 5, 32
    a = herm_cond(n, kappa, seed=21)
    rng = np.random.default_rng(22)
    b = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    x, it, info = ora.zcposv(uplo, a, b, nb)
    assert info == 0
    assert 0 <= it <= 6, it
    eps = 2.0 ** -53
    res = (np.abs(b - a @ x).sum(1).max()
           / (n * np.abs(a).sum(1).max() * np.abs(x).sum(1).max() * eps))
    assert res < 10, res
    want = np.linalg.solve(a, b)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 100 * kappa * eps


def test_zcposv_real_input_matches_dsposv():
    """Real data as complex: same iteration count as the real solver."""
    n, nrhs, nb = 128, 4, 32
    rng = np.random.default_rng(31)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q * np.logspace(0, -5, n)) @ q.T
    a = (a + a.T) / 2
    b = rng.standard_normal((n, nrhs))
    xz, itz, infoz = ora.zcposv(L, a.astype(np.complex128),
                                b.astype(np.complex128), nb)
    xd, itd, infod = ora.dsposv(L, a, b, nb)
    assert infoz == 0 and infod == 0
    assert itz == itd
    assert np.abs(xz.imag).max() == 0
    # both are FP64-accurate solutions (cond 1e5): they agree to cond * eps,
    # their residual products summing in different orders
    assert np.linalg.norm(xz.real - xd) / np.linalg.norm(xd) < 1e5 * 2.0 ** -50


def test_zcposv_fallback_and_breakdown():
    """cond 1e8: refinement from the single complex factor cannot reach the
    FP64 test in 30 steps, iter = -31 and the double complex factorization
    solves the system; cond 1e9: the single complex factorization breaks
    down and its info fails the call (reference zcposv.c:452-460,
    :524-538)."""
    n, nrhs, nb = 96, 2, 32
    rng = np.random.default_rng(42)
    b = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    a = herm_cond(n, 1e8, seed=41)
    x, it, info = ora.zcposv(L, a, b, nb)
    assert info == 0 and it == -31
    want = np.linalg.solve(a, b)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e8 * 2.0 ** -45
    a = herm_cond(n, 1e9, seed=41)
    _, _, info = ora.zcposv(L, a, b, nb)
    assert info > 0


def test_zcposv_exact_fallback_and_breakdown():
    """Gaussian-integer factors make the single complex factorization exact:
    A = A0 + e diag(A0) below single resolution diverges under refinement
    (-31, then the double complex solve); A1 singular in single precision
    breaks down at the zero column (info 10)."""
    n = 16
    rng = np.random.default_rng(3)
    p = np.array([1, 1j, -1, -1j])[rng.integers(0, 4, n)]
    t = np.eye(n) - np.tril(np.ones((n, n)), -1)
    l0 = (p[:, None] * t) * p.conj()[None, :]
    a0 = l0 @ l0.conj().T
    d = np.diag(np.diag(a0).real)
    rho = max(abs(np.linalg.eigvals(np.linalg.solve(a0, d))))
    a = a0 + (1.5 / rho) * d
    b = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    for uplo in (L, U):
        x, it, info = ora.zcposv(uplo, a, b, 4)
        assert info == 0 and it == -31
        res = (np.abs(b - a @ x).sum(1).max()
               / (n * np.abs(a).sum(1).max() * np.abs(x).sum(1).max()
                  * 2.0 ** -53))
        assert res < 10
    l0[9, 9] = 0
    a1 = l0 @ l0.conj().T + 1e-9 * np.eye(n)
    for uplo in (L, U):
        assert ora.zcposv(uplo, a1, b, 8)[2] == 10


def test_zlag2c_semantics():
    rng = np.random.default_rng(51)
    a = rng.standard_normal((9, 4)) + 1j * rng.standard_normal((9, 4))
    s, info = ora.zlag2c(a)
    assert info == 0
    assert np.array_equal(s, a.astype(np.complex64))
    back = ora.clag2z(s)
    assert np.array_equal(back, s.astype(np.complex128))
    big = 2.0 * float(np.finfo(np.float32).max)
    for v in (big + 0j, 1j * big, -big + 1j, 1 - 1j * big):
        b = a.copy()
        b[3, 2] = v
        assert ora.zlag2c(b)[1] == 1
    # the largest finite single value converts
    b = a.copy()
    b[0, 0] = float(np.finfo(np.float32).max) * (1 - 1j)
    assert ora.zlag2c(b)[1] == 0


@pytest.mark.parametrize("uplo", [L, U])
def test_zlanhe_and_dzamax(uplo):
    n = 45
    a = herm_pd(n, seed=61)
    full = a.copy()
    stored = a.copy()
    # the unstored triangle holds garbage and the diagonal an imaginary part
    if uplo == L:
        stored[np.triu_indices(n, 1)] = 1e3
    else:
        stored[np.tril_indices(n, -1)] = 1e3
    stored[np.diag_indices(n)] += 0.5j
    want = np.abs(full).sum(1).max()
    assert abs(ora.zlanhe_inf(uplo, stored) - want) <= 1e-13 * want
    x = a[:, :6]
    want = (np.abs(x.real) + np.abs(x.imag)).max(0)
    assert np.array_equal(ora.dzamax(x), want)
