"""GPU parity tests of the complex (Hermitian) Cholesky path through the C
ABI: plasma_{z,c}potrf / potrs / posv, plasma_zcposv, plasma_zlag2c /
clag2z and the tile forms, against the oracle's complex restatement
(pinned to LAPACK zpotrf / cpotrf / zpotrs in test_oracle_complex.py).

Checks (the reference tester's, test/test_zpotrf.c:124-152,
test_zposv.c:144-168, test_zcposv.c:327-373):
  * potrf: ||L - L_oracle||_F / ||L_oracle||_F <= 1e-12 (z) or 1e-5 (c),
    a real diagonal, the other triangle and padding rows untouched;
  * posv / potrs: X against the oracle and the scaled residual
    ||B - AX||_inf / (n ||A||_inf ||X||_inf eps) < 10;
  * zcposv: iteration count within one of the oracle's, X within
    cond * eps of it, the residual test, the -31 fallback and a factor
    breakdown's info;
  * conversions bit for bit, overflow of either part giving 1."""
import ctypes

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper
EPS = {np.complex128: 2.0 ** -53, np.complex64: 2.0 ** -24}


def herm_pd(n, seed=0, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = g @ g.conj().T / n + np.eye(n)
    return a.astype(dtype)


def herm_cond(n, kappa, seed=0):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q, _ = np.linalg.qr(g)
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    a = (q * s) @ q.conj().T
    return (a + a.conj().T) / 2


def crhs(n, nrhs, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    return np.asfortranarray(b.astype(dtype))


def tri(a, lower):
    return np.tril(a) if lower else np.triu(a)


def ferr(f, ref, lower):
    return (np.linalg.norm(tri(f, lower) - tri(ref, lower))
            / np.linalg.norm(tri(ref, lower)))


def scaled_residual(a, x, b, eps):
    n = a.shape[0]
    r = b - a @ x
    return (np.abs(r).sum(1).max()
            / (n * np.abs(a).sum(1).max() * np.abs(x).sum(1).max() * eps))


def _pfx(dtype):
    return "z" if dtype == np.complex128 else "c"


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb", [(1, 16), (37, 16), (200, 64), (333, 128),
                                  (520, 96), (700, 256)])
def test_potrf(plasma, nb_guard, dtype, uplo, n, nb):
    pa.set_param(pa.PlasmaNb, nb)
    lda = n + 2
    A = np.zeros((lda, n), dtype=dtype, order="F")
    A[:n, :n] = herm_pd(n, seed=n, dtype=dtype)
    A[n:, :] = 7 - 3j
    # the diagonal's imaginary part is not read
    A[np.arange(n), np.arange(n)] += 0.25j
    a = A.copy(order="F")
    info = pa.potrf(uplo, n, a, lda, prec=_pfx(dtype))
    assert info == 0
    ref, rinfo = ora.zpotrf(uplo, A[:n, :n], nb, dtype=dtype)
    assert rinfo == 0
    lo = uplo == L
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    assert ferr(a[:n, :n], ref, lo) <= tol
    assert np.all(np.diag(a[:n, :n]).imag == 0)
    other = np.triu(A[:n, :n], 1) if lo else np.tril(A[:n, :n], -1)
    got_other = np.triu(a[:n, :n], 1) if lo else np.tril(a[:n, :n], -1)
    assert np.array_equal(got_other, other)
    assert np.array_equal(a[n:, :], A[n:, :])


@pytest.mark.parametrize("uplo", [L, Up])
def test_zpotrf_device_large(plasma, nb_guard, uplo):
    """A device-resident order-2048 factor with the default tile size."""
    import torch
    import scipy.linalg.lapack as la
    n = 2048
    pa.set_param(pa.PlasmaNb, 128)
    A = herm_pd(n, seed=77)
    # column-major on the device: tensor row j holds column j
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    torch.cuda.synchronize()
    assert pa.potrf(uplo, n, dA, n, prec="z") == 0
    f = dA.cpu().numpy().T
    ref, info = la.zpotrf(A, lower=uplo == L, clean=False)
    assert info == 0
    assert ferr(f, ref, uplo == L) <= 1e-12


@pytest.mark.parametrize("uplo", [L, Up])
def test_not_positive_definite(plasma, nb_guard, uplo):
    n, nb, j0 = 300, 64, 181
    pa.set_param(pa.PlasmaNb, nb)
    A = herm_pd(n, seed=5)
    A[j0, j0] = -3.0
    a = np.asfortranarray(A.copy())
    info = pa.potrf(uplo, n, a, n, prec="z")
    _, rinfo = ora.zpotrf(uplo, A, nb)
    assert info == rinfo == j0 + 1
    # a failed factorization leaves pA as it was
    assert np.array_equal(a, A)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nrhs,nb", [(150, 1, 64), (333, 40, 128),
                                       (260, 150, 64)])
def test_posv_potrs(plasma, nb_guard, dtype, uplo, n, nrhs, nb):
    pa.set_param(pa.PlasmaNb, nb)
    A = herm_pd(n, seed=n + nrhs, dtype=dtype)
    B = crhs(n, nrhs, seed=3, dtype=dtype)
    p = _pfx(dtype)
    a = np.asfortranarray(A.copy())
    x = B.copy(order="F")
    assert pa.posv(uplo, n, nrhs, a, n, x, n, prec=p) == 0
    fo, xo, info = ora.zposv(uplo, A, B, nb, dtype=dtype)
    assert info == 0
    eps = EPS[dtype]
    tol = 1e-12 if dtype == np.complex128 else 1e-4
    assert ferr(a, fo, uplo == L) <= tol
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 100 * tol
    A64, B64, x64 = (v.astype(np.complex128) for v in (A, B, x))
    assert scaled_residual(A64, x64, B64, eps) < 10
    # potrs from the oracle's factor
    x2 = B.copy(order="F")
    f = np.asfortranarray(fo)
    assert pa.potrs(uplo, n, nrhs, f, n, x2, n, prec=p) == 0
    assert np.linalg.norm(x2 - xo) / np.linalg.norm(xo) <= 100 * tol


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_zcposv(plasma, nb_guard, uplo, kappa):
    n, nrhs, nb = 600, 8, 128
    pa.set_param(pa.PlasmaNb, nb)
    A = np.asfortranarray(herm_cond(n, kappa, seed=int(np.log10(kappa))))
    B = crhs(n, nrhs, seed=9)
    X = np.zeros_like(B, order="F")
    a0 = A.copy()
    rc, it = pa.zcposv(uplo, n, nrhs, A, n, B, n, X, n)
    assert rc == 0
    Xo, ito, info = ora.zcposv(uplo, A, B, nb)
    assert info == 0
    assert abs(it - ito) <= 1 and it >= 0, (it, ito)
    eps = 2.0 ** -53
    assert scaled_residual(A, X, B, eps) < 10
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 100 * kappa * eps
    assert np.array_equal(A, a0)


def exact_factor(n, seed):
    """L0 = P (I - strictly lower ones) P^H with P a diagonal of units
    (1, i, -1, -i): a Gaussian-integer factor, so A0 = L0 L0^H and its
    single complex factorization are exact in any summation order."""
    rng = np.random.default_rng(seed)
    p = np.array([1, 1j, -1, -1j])[rng.integers(0, 4, n)]
    t = np.eye(n) - np.tril(np.ones((n, n)), -1)
    l0 = (p[:, None] * t) * p.conj()[None, :]
    return l0, l0 @ l0.conj().T


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("nb", [4, 8])
def test_zcposv_fallback(plasma, nb_guard, uplo, nb):
    """A = A0 + e D (D = diag(A0)) with e below single precision's
    resolution: the single factor is A0's, exactly, and refinement with it
    diverges (spectral radius 1.5 of A0^{-1} e D), so after 30 steps
    iter = -31 and the double complex factorization solves (reference
    zcposv.c:524-538), on the GPU as in the oracle."""
    n, nrhs = 16, 2
    pa.set_param(pa.PlasmaNb, nb)
    _, a0 = exact_factor(n, 3)
    d = np.diag(np.diag(a0).real)
    rho = max(abs(np.linalg.eigvals(np.linalg.solve(a0, d))))
    A = np.asfortranarray(a0 + (1.5 / rho) * d)
    assert np.array_equal(A.astype(np.complex64), a0.astype(np.complex64))
    B = crhs(n, nrhs, seed=5)
    X = np.zeros_like(B, order="F")
    rc, it = pa.zcposv(uplo, n, nrhs, A, n, B, n, X, n)
    _, ito, info = ora.zcposv(uplo, A, B, nb)
    assert rc == 0 and info == 0
    assert it == ito == -31
    assert scaled_residual(A, X, B, 2.0 ** -53) < 10


@pytest.mark.parametrize("uplo", [L, Up])
def test_zcposv_single_factor_breakdown(plasma, nb_guard, uplo):
    """A = A1 + 1e-9 I with A1 singular and exact in single precision (its
    factor has a zero at column 9): the single complex factorization meets
    a zero pivot at column 10 and the call fails with that info, leaving pX
    as it was (reference zcposv.c:452-460)."""
    n, nrhs = 16, 2
    pa.set_param(pa.PlasmaNb, 8)
    l1, _ = exact_factor(n, 3)
    l1[9, 9] = 0
    A = np.asfortranarray(l1 @ l1.conj().T + 1e-9 * np.eye(n))
    B = crhs(n, nrhs, seed=5)
    X = np.full_like(B, 7 + 7j, order="F")
    rc, _ = pa.zcposv(uplo, n, nrhs, A, n, B, n, X, n)
    _, _, info = ora.zcposv(uplo, A, B, 8)
    assert rc == info == 10
    assert np.all(X == 7 + 7j)


def test_lag2_conversions(plasma, nb_guard):
    lib = pa.load()
    pa.set_param(pa.PlasmaNb, 64)
    m, n = 130, 70
    rng = np.random.default_rng(12)
    A = np.asfortranarray(rng.standard_normal((m, n))
                          + 1j * rng.standard_normal((m, n)))
    S = np.zeros((m, n), dtype=np.complex64, order="F")
    assert lib.plasma_zlag2c(m, n, A.ctypes.data, m, S.ctypes.data, m) == 0
    So, info = ora.zlag2c(A)
    assert info == 0 and np.array_equal(S, So)
    Z = np.zeros((m, n), dtype=np.complex128, order="F")
    assert lib.plasma_clag2z(m, n, S.ctypes.data, m, Z.ctypes.data, m) == 0
    assert np.array_equal(Z, ora.clag2z(S))
    big = 2.0 * float(np.finfo(np.float32).max)
    for v in (big, 1j * big):
        A2 = A.copy(order="F")
        A2[100, 60] = v
        assert lib.plasma_zlag2c(m, n, A2.ctypes.data, m, S.ctypes.data,
                                 m) == 1


def test_argument_codes(plasma, nb_guard):
    lib = pa.load()
    a = np.zeros((4, 4), dtype=np.complex128, order="F")
    p = a.ctypes.data
    assert lib.plasma_zpotrf(0, 4, p, 4) == -1
    assert lib.plasma_zpotrf(L, -1, p, 4) == -2
    assert lib.plasma_zpotrf(L, 4, p, 3) == -4
    assert lib.plasma_cposv(L, 4, 1, p, 4, p, 3) == -7
    it = ctypes.c_int(0)
    assert lib.plasma_zcposv(L, 4, 1, p, 4, p, 4, p, 3, ctypes.byref(it)) == -9


def test_omp_tile_forms_general_descriptors(plasma, nb_guard):
    """plasma_omp_zpotrf / zpotrs on caller-created general descriptors
    (the reference's zpotrs uses a general A, compute/zpotrs.c:127-137),
    filled by plasma_omp_zge2desc, synchronised by plasma_sequence_sync."""
    lib = pa.load()
    n, nrhs, nb = 400, 5, 128
    A = np.asfortranarray(herm_pd(n, seed=31))
    B = crhs(n, nrhs, seed=32)
    dA, dB = pa.plasma_desc_t(), pa.plasma_desc_t()
    Z = pa.PlasmaComplexDouble
    assert lib.plasma_desc_general_create(Z, nb, nb, n, n, 0, 0, n, n,
                                          ctypes.byref(dA)) == 0
    assert lib.plasma_desc_general_create(Z, nb, nb, n, nrhs, 0, 0, n, nrhs,
                                          ctypes.byref(dB)) == 0
    s, r = pa.plasma_sequence_t(), pa.plasma_request_t()
    lib.plasma_sequence_init(ctypes.byref(s))
    lib.plasma_request_init(ctypes.byref(r))
    S, R = ctypes.byref(s), ctypes.byref(r)
    F = np.zeros((n, n), dtype=np.complex128, order="F")
    X = np.zeros((n, nrhs), dtype=np.complex128, order="F")
    lib.plasma_omp_zge2desc(A.ctypes.data, n, dA, S, R)
    lib.plasma_omp_zge2desc(B.ctypes.data, n, dB, S, R)
    lib.plasma_omp_zpotrf(L, dA, S, R)
    lib.plasma_omp_zpotrs(L, dA, dB, S, R)
    lib.plasma_omp_zdesc2ge(dA, F.ctypes.data, n, S, R)
    lib.plasma_omp_zdesc2ge(dB, X.ctypes.data, n, S, R)
    assert lib.plasma_sequence_sync(S) == 0
    fo, xo, info = ora.zposv(L, A, B, nb)
    assert info == 0
    assert ferr(F, fo, True) <= 1e-12
    assert np.linalg.norm(X - xo) / np.linalg.norm(xo) <= 1e-11
    lib.plasma_desc_destroy(ctypes.byref(dA))
    lib.plasma_desc_destroy(ctypes.byref(dB))
