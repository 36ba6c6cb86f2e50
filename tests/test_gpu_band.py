"""GPU parity tests of the band Cholesky through the C ABI:
plasma_{d,s}pbtrf, plasma_{d,s}pbtrs, plasma_{d,s}pbsv and the tile forms
(plasma_omp_dpbtrf / dpbtrs / dpbsv on caller-created band descriptors with
plasma_omp_dpb2desc / ddesc2pb), against the oracle's restatement of
compute/pzpbtrf.c and pztbsm.c (pinned to LAPACK in test_oracle_band.py).

Checks: the factor against the oracle's (FP64 within 1e-12 relative, FP32
within 1e-5), the tester's scaled residual of pbsv / pbtrs, LAPACK's info
for a leading minor that is not positive definite with AB left unchanged,
band widths below, at and above nb with n not a multiple of nb, both
triangles, host and device arrays, padding rows of AB passing through, and
the reference's argument codes (compute/zpbtrf.c:91-110, zpbtrs.c:94-118,
zpbsv.c:105-129)."""
import ctypes

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _band as BD
from tests import _util as U

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper
OU = {L: ora.LOWER, Up: ora.UPPER}


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def _host(t):
    return np.asfortranarray(t.cpu().numpy().T)


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,kd,nb", [(300, 0, 64), (300, 5, 64),
                                     (517, 63, 64), (517, 64, 64),
                                     (700, 200, 128), (1000, 131, 256),
                                     (640, 639, 128)])
@pytest.mark.parametrize("device", [False, True])
def test_dpbtrf(plasma, nb_guard, uplo, n, kd, nb, device):
    import torch
    pa.set_param(pa.PlasmaNb, nb)
    A = BD.spd_band(n, kd, seed=n + kd)
    ab = BD.to_band(A, kd, uplo == L, ldab=kd + 2)
    ab[kd + 1:, :] = 7.0                     # padding row: passes through
    lib = pa.load()
    if device:
        d = _dev(ab)
        torch.cuda.synchronize()
        assert lib.plasma_dpbtrf(uplo, n, kd, d.data_ptr(), kd + 2) == 0
        got = _host(d)
    else:
        got = ab.copy(order="F")
        assert lib.plasma_dpbtrf(uplo, n, kd, got.ctypes.data, kd + 2) == 0
    ora.set_blas(ora.MKL)
    ref, info = ora.pbtrf(OU[uplo], ab, kd, nb)
    assert info == 0
    band = slice(0, kd + 1)
    assert np.linalg.norm(got[band] - ref[band]) <= 1e-12 * np.linalg.norm(
        ref[band])
    assert np.all(got[kd + 1:] == 7.0)


@pytest.mark.parametrize("uplo", [L, Up])
def test_spbtrf_and_spbsv(plasma, nb_guard, uplo):
    pa.set_param(pa.PlasmaNb, 128)
    n, kd = 900, 150
    A = BD.spd_band(n, kd, seed=5)
    ab = BD.to_band(A, kd, uplo == L).astype(np.float32, order="F")
    lib = pa.load()
    got = ab.copy(order="F")
    assert lib.plasma_spbtrf(uplo, n, kd, got.ctypes.data, kd + 1) == 0
    ora.set_blas(ora.MKL)
    ref, info = ora.pbtrf(OU[uplo], ab, kd, 128, dtype=np.float32)
    assert info == 0
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    B = U.rhs(n, 3, seed=6, dtype=np.float32)
    a2, b = ab.copy(order="F"), B.copy(order="F")
    assert lib.plasma_spbsv(uplo, n, kd, 3, a2.ctypes.data, kd + 1,
                            b.ctypes.data, n) == 0
    res = U.scaled_residual(A.astype(np.float32).astype(np.float64),
                            b.astype(np.float64), B.astype(np.float64))
    assert res < 10 * U.EPS32


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,kd,nb,nrhs", [(800, 70, 128, 9),
                                          (333, 129, 64, 40)])
def test_dpbsv_and_dpbtrs(plasma, nb_guard, uplo, n, kd, nb, nrhs):
    import torch
    pa.set_param(pa.PlasmaNb, nb)
    A = BD.spd_band(n, kd, seed=9)
    ab = BD.to_band(A, kd, uplo == L)
    B = U.rhs(n, nrhs, seed=10)
    lib = pa.load()
    dab, dB = _dev(ab), _dev(B)
    torch.cuda.synchronize()
    assert lib.plasma_dpbsv(uplo, n, kd, nrhs, dab.data_ptr(), kd + 1,
                            dB.data_ptr(), n) == 0
    F, X = _host(dab), _host(dB)
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    ora.set_blas(ora.MKL)
    fref, Xref, info = ora.pbsv(OU[uplo], ab, kd, B, nb)
    assert info == 0
    assert np.linalg.norm(F - fref) <= 1e-12 * np.linalg.norm(fref)
    assert np.linalg.norm(X - Xref) <= 1e-11 * np.linalg.norm(Xref)
    # pbtrs with the factor: host arrays
    b2 = B.copy(order="F")
    f2 = F.copy(order="F")
    assert lib.plasma_dpbtrs(uplo, n, kd, nrhs, f2.ctypes.data, kd + 1,
                             b2.ctypes.data, n) == 0
    assert np.array_equal(f2, F)                      # factor not modified
    assert U.scaled_residual(A, b2, B) < 10 * U.EPS64


@pytest.mark.parametrize("uplo", [L, Up])
def test_dpbtrf_not_pd(plasma, nb_guard, uplo):
    pa.set_param(pa.PlasmaNb, 64)
    n, kd = 400, 30
    A = BD.spd_band(n, kd, seed=11)
    A[233, 233] = -100.0
    ab = BD.to_band(A, kd, uplo == L)
    got = ab.copy(order="F")
    info = pa.load().plasma_dpbtrf(uplo, n, kd, got.ctypes.data, kd + 1)
    ora.set_blas(ora.MKL)
    _, rinfo = ora.pbtrf(OU[uplo], ab, kd, 64)
    assert info == rinfo > 0
    assert np.array_equal(got, ab)


def test_band_tile_forms(plasma, nb_guard):
    """plasma_omp_dpbsv on a caller-created band descriptor, with the band
    array translated in and out by plasma_omp_dpb2desc / ddesc2pb, in one
    sequence; then plasma_omp_dpbtrs with the factor."""
    lib = pa.load()
    n, kd, nb, nrhs = 600, 90, 128, 4
    A = BD.spd_band(n, kd, seed=12)
    ab = BD.to_band(A, kd, True)
    B = U.rhs(n, nrhs, seed=13)
    lm = nb * (1 + (kd + nb - 1) // nb)
    dAB, dB = pa.plasma_desc_t(), pa.plasma_desc_t()
    assert lib.plasma_desc_general_band_create(
        pa.PlasmaRealDouble, L, nb, nb, lm, n, 0, 0, n, n, kd, kd,
        ctypes.byref(dAB)) == 0
    assert lib.plasma_desc_general_create(pa.PlasmaRealDouble, nb, nb, n,
                                          nrhs, 0, 0, n, nrhs,
                                          ctypes.byref(dB)) == 0
    s, r = pa.plasma_sequence_t(), pa.plasma_request_t()
    lib.plasma_sequence_init(ctypes.byref(s))
    lib.plasma_request_init(ctypes.byref(r))
    S, R = ctypes.byref(s), ctypes.byref(r)
    fab = ab.copy(order="F")
    X = B.copy(order="F")
    lib.plasma_omp_dpb2desc(fab.ctypes.data, kd + 1, dAB, S, R)
    lib.plasma_omp_dge2desc(X.ctypes.data, n, dB, S, R)
    lib.plasma_omp_dpbsv(L, dAB, dB, S, R)
    lib.plasma_omp_ddesc2pb(dAB, fab.ctypes.data, kd + 1, S, R)
    lib.plasma_omp_ddesc2ge(dB, X.ctypes.data, n, S, R)
    assert lib.plasma_sequence_sync(S) == 0
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    ora.set_blas(ora.MKL)
    fref, _ = ora.pbtrf(ora.LOWER, ab, kd, nb)
    assert np.linalg.norm(fab - fref) <= 1e-12 * np.linalg.norm(fref)
    # solve again with the factor still in the descriptor
    X2 = B.copy(order="F")
    lib.plasma_omp_dge2desc(X2.ctypes.data, n, dB, S, R)
    lib.plasma_omp_dpbtrs(L, dAB, dB, S, R)
    lib.plasma_omp_ddesc2ge(dB, X2.ctypes.data, n, S, R)
    assert lib.plasma_sequence_sync(S) == 0
    assert np.abs(X2 - X).max() <= 1e-12 * np.abs(X).max()
    # a general descriptor is not a band one
    lib.plasma_omp_dpbtrf(L, dB, S, R)
    assert lib.plasma_sequence_sync(S) == pa.PlasmaErrorIllegalValue
    lib.plasma_desc_destroy(ctypes.byref(dAB))
    lib.plasma_desc_destroy(ctypes.byref(dB))


def test_band_argument_codes(plasma, nb_guard):
    lib = pa.load()
    a = np.zeros((4, 10), order="F")
    p = a.ctypes.data
    assert lib.plasma_dpbtrf(7, 10, 3, p, 4) == -1
    assert lib.plasma_dpbtrf(L, -1, 3, p, 4) == -2
    assert lib.plasma_dpbtrf(L, 10, -1, p, 4) == -3
    assert lib.plasma_dpbtrf(L, 10, 3, p, 3) == -5
    assert lib.plasma_dpbtrf(L, 0, 3, p, 4) == 0
    assert lib.plasma_dpbtrs(7, 10, 3, 1, p, 4, p, 10) == -1
    assert lib.plasma_dpbtrs(L, -1, 3, 1, p, 4, p, 10) == -2
    assert lib.plasma_dpbtrs(L, 10, -1, 1, p, 4, p, 10) == -4
    assert lib.plasma_dpbtrs(L, 10, 3, -1, p, 4, p, 10) == -5
    assert lib.plasma_dpbtrs(L, 10, 3, 1, p, 3, p, 10) == -7
    assert lib.plasma_dpbtrs(L, 10, 3, 1, p, 4, p, 9) == -10
    assert lib.plasma_dpbsv(7, 10, 3, 1, p, 4, p, 10) == -1
    assert lib.plasma_dpbsv(L, -1, 3, 1, p, 4, p, 10) == -2
    assert lib.plasma_dpbsv(L, 10, -1, 1, p, 4, p, 10) == -3
    assert lib.plasma_dpbsv(L, 10, 3, -1, p, 4, p, 10) == -4
    assert lib.plasma_dpbsv(L, 10, 3, 1, p, 3, p, 10) == -6
    assert lib.plasma_dpbsv(L, 10, 3, 1, p, 4, p, 9) == -8
    assert lib.plasma_dpbsv(L, 10, 3, 0, p, 4, p, 10) == 0
