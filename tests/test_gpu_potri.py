"""GPU parity tests of the inverse from the Cholesky factor through the C
ABI: plasma_dpotri / plasma_dpoinv (+ s, + the tile forms) against the
oracle's restatement of compute/zpotri.c / zpoinv.c (pinned to LAPACK's
dpotri in test_oracle_potri.py).

Checks: the inverse's triangle against the oracle (FP64 within 1e-12
relative), A A^{-1} = I, the other triangle untouched, LAPACK's info for a
zero on the factor's diagonal (and potrf's info for poinv of a matrix that
is not positive definite, pA unchanged), host and device pointers, both
triangles, edge tiles, the argument codes."""
import ctypes

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper


def _factor(A, lower):
    F = np.linalg.cholesky(A)
    return np.asfortranarray(np.tril(F) if lower else np.triu(F.T))


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb", [(1, 64), (300, 128), (1000, 256),
                                  (1100, 128)])
def test_dpotri(plasma, nb_guard, uplo, n, nb):
    pa.set_param(pa.PlasmaNb, nb)
    lower = uplo == L
    A = U.ggt_spd(n, seed=n)
    F = _factor(A, lower)
    # garbage in the other triangle: must pass through unchanged
    junk = np.triu(np.full((n, n), 3.25), 1) if lower else \
        np.tril(np.full((n, n), 3.25), -1)
    F = np.asfortranarray(F + junk)
    a = F.copy(order="F")
    assert pa.load().plasma_dpotri(uplo, n, a.ctypes.data, n) == 0
    ref, info = ora.potri(ora.LOWER if lower else ora.UPPER, F, nb)
    assert info == 0
    tri = np.tril if lower else np.triu
    assert np.linalg.norm(tri(a) - tri(ref)) <= 1e-12 * np.linalg.norm(
        tri(ref))
    assert np.array_equal(a - tri(a), junk)
    full = U.sym_full(a, lower)
    assert np.abs(A @ full - np.eye(n)).max() <= 1e-10


@pytest.mark.parametrize("uplo", [L, Up])
def test_dpoinv_device(plasma, nb_guard, uplo):
    import torch
    n, nb = 900, 128
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=9)
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    torch.cuda.synchronize()
    assert pa.load().plasma_dpoinv(uplo, n, dA.data_ptr(), n) == 0
    a = dA.cpu().numpy().T
    lower = uplo == L
    ref, info = ora.poinv(ora.LOWER if lower else ora.UPPER, A, nb)
    assert info == 0
    tri = np.tril if lower else np.triu
    assert np.linalg.norm(tri(a) - tri(ref)) <= 1e-12 * np.linalg.norm(
        tri(ref))
    other = np.triu(A, 1) if lower else np.tril(A, -1)
    assert np.array_equal(a - tri(a), other)


def test_spotri_spoinv(plasma, nb_guard):
    n, nb = 500, 128
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=2)
    ref = np.linalg.inv(A)
    a = np.asfortranarray(A.astype(np.float32))
    assert pa.load().plasma_spoinv(L, n, a.ctypes.data, n) == 0
    err = np.abs(np.tril(a) - np.tril(ref)).max() / np.abs(ref).max()
    assert err <= 1e-5
    F = np.asfortranarray(np.linalg.cholesky(A).astype(np.float32))
    assert pa.load().plasma_spotri(L, n, F.ctypes.data, n) == 0
    err = np.abs(np.tril(F) - np.tril(ref)).max() / np.abs(ref).max()
    assert err <= 1e-5


@pytest.mark.parametrize("zero", [0, 130, 299])
def test_dpotri_zero_diagonal_info(plasma, nb_guard, zero):
    n, nb = 300, 128
    pa.set_param(pa.PlasmaNb, nb)
    F = _factor(U.ggt_spd(n, seed=4), True)
    F[zero, zero] = 0.0
    a = F.copy(order="F")
    info = pa.load().plasma_dpotri(L, n, a.ctypes.data, n)
    _, oinfo = ora.potri(ora.LOWER, F, nb)
    assert info == oinfo == zero + 1
    assert np.array_equal(a, F)          # failed: pA not written back


def test_dpoinv_not_pd(plasma, nb_guard):
    n, nb = 300, 128
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=5)
    A[:, 200] = 0.0
    A[200, :] = 0.0
    a = A.copy(order="F")
    assert pa.load().plasma_dpoinv(L, n, a.ctypes.data, n) == 201
    assert np.array_equal(a, A)


def test_potri_tile_form_and_codes(plasma, nb_guard):
    lib = pa.load()
    n, nb = 400, 128
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=6)
    F = _factor(A, True)
    D = pa.plasma_desc_t()
    assert lib.plasma_desc_general_create(pa.PlasmaRealDouble, nb, nb, n, n,
                                          0, 0, n, n, ctypes.byref(D)) == 0
    seq, req = pa.plasma_sequence_t(), pa.plasma_request_t()
    lib.plasma_sequence_init(ctypes.byref(seq))
    lib.plasma_request_init(ctypes.byref(req))
    out = np.zeros((n, n), order="F")
    lib.plasma_omp_dge2desc(F.ctypes.data, n, D, ctypes.byref(seq),
                            ctypes.byref(req))
    lib.plasma_omp_dpotri(L, D, ctypes.byref(seq), ctypes.byref(req))
    lib.plasma_omp_ddesc2ge(D, out.ctypes.data, n, ctypes.byref(seq),
                            ctypes.byref(req))
    assert lib.plasma_sequence_sync(ctypes.byref(seq)) == 0
    ref = np.linalg.inv(A)
    assert np.abs(np.tril(out) - np.tril(ref)).max() <= \
        1e-12 * np.abs(ref).max() * 100
    lib.plasma_omp_dpotri(12345, D, ctypes.byref(seq), ctypes.byref(req))
    assert seq.status == pa.PlasmaErrorIllegalValue
    lib.plasma_desc_destroy(ctypes.byref(D))
    a = np.eye(4, order="F")
    assert lib.plasma_dpotri(999, 4, a.ctypes.data, 4) == -1
    assert lib.plasma_dpotri(L, -1, a.ctypes.data, 4) == -2
    assert lib.plasma_dpotri(L, 4, a.ctypes.data, 3) == -4
    assert lib.plasma_dpoinv(L, 0, a.ctypes.data, 4) == 0
