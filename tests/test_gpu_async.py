"""GPU tests of the asynchronous tile interface through the C ABI, with
descriptors the caller creates and fills, as a PLASMA application uses it:

  * plasma_omp_dpotrs with a *general* descriptor for A (the reference's
    plasma_dpotrs translates A into a general descriptor, compute/zpotrs.c:
    127-137), after plasma_omp_dpotrf on the same general descriptor;
  * plasma_omp_dsposv with general A, As and Xs (compute/zcposv.c:171-224),
    plus plasma_omp_dlag2s / plasma_omp_slag2d on caller descriptors;
  * more than 256 sequences in flight, one of them failing: the library
    reuses its device status words only after folding each pending failure
    into its sequence, so the failing sequence still reports LAPACK's info.
Results against the oracle (pinned to LAPACK in test_oracle.py)."""
import ctypes

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L = pa.PlasmaLower


def _seq(lib):
    s, r = pa.plasma_sequence_t(), pa.plasma_request_t()
    lib.plasma_sequence_init(ctypes.byref(s))
    lib.plasma_request_init(ctypes.byref(r))
    return s, r


def _general(lib, prec, nb, m, n):
    d = pa.plasma_desc_t()
    assert lib.plasma_desc_general_create(prec, nb, nb, m, n, 0, 0, m, n,
                                          ctypes.byref(d)) == 0
    return d


def test_omp_potrf_potrs_general_descriptor(plasma, nb_guard):
    lib = pa.load()
    n, nrhs, nb = 700, 9, 128
    A = U.ggt_spd(n, seed=21)
    B = U.rhs(n, nrhs, seed=22)
    dA = _general(lib, pa.PlasmaRealDouble, nb, n, n)
    dB = _general(lib, pa.PlasmaRealDouble, nb, n, nrhs)
    s, r = _seq(lib)
    S, R = ctypes.byref(s), ctypes.byref(r)
    F = np.zeros((n, n), order="F")
    X = np.zeros((n, nrhs), order="F")
    lib.plasma_omp_dge2desc(A.ctypes.data, n, dA, S, R)
    lib.plasma_omp_dge2desc(B.ctypes.data, n, dB, S, R)
    lib.plasma_omp_dpotrf(L, dA, S, R)
    lib.plasma_omp_dpotrs(L, dA, dB, S, R)
    lib.plasma_omp_ddesc2ge(dA, F.ctypes.data, n, S, R)
    lib.plasma_omp_ddesc2ge(dB, X.ctypes.data, n, S, R)
    assert lib.plasma_sequence_sync(S) == 0
    ref, info = ora.potrf(ora.LOWER, A, nb)
    assert info == 0
    assert U.factor_error(F, ref, True) <= 1e-12
    Xo, info = ora.potrs(ora.LOWER, ref, B, nb)
    assert np.abs(X - Xo).max() <= 1e-12 * np.abs(Xo).max()
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    lib.plasma_desc_destroy(ctypes.byref(dA))
    lib.plasma_desc_destroy(ctypes.byref(dB))


def test_omp_dsposv_general_descriptors(plasma, nb_guard):
    lib = pa.load()
    n, nrhs, nb = 512, 6, 128
    A = U.householder_spd(n, 1e4, seed=3)
    B = U.rhs(n, nrhs, seed=4)
    D, Sg = pa.PlasmaRealDouble, pa.PlasmaRealFloat
    dA, dB, dX, dR = (_general(lib, D, nb, n, n), _general(lib, D, nb, n, nrhs),
                      _general(lib, D, nb, n, nrhs),
                      _general(lib, D, nb, n, nrhs))
    dAs, dXs = _general(lib, Sg, nb, n, n), _general(lib, Sg, nb, n, nrhs)
    s, r = _seq(lib)
    S, R = ctypes.byref(s), ctypes.byref(r)
    it = ctypes.c_int(-99)
    X = np.zeros((n, nrhs), order="F")
    lib.plasma_omp_dge2desc(A.ctypes.data, n, dA, S, R)
    lib.plasma_omp_dge2desc(B.ctypes.data, n, dB, S, R)
    lib.plasma_omp_dsposv(L, dA, dB, dX, dAs, dXs, dR, None, None, None,
                          ctypes.byref(it), S, R)
    lib.plasma_omp_ddesc2ge(dX, X.ctypes.data, n, S, R)
    assert lib.plasma_sequence_sync(S) == 0
    Xo, ito, info = ora.dsposv(ora.LOWER, A, B, nb)
    assert info == 0
    assert abs(it.value - ito) <= 1, (it.value, ito)
    assert np.linalg.norm(X - Xo) <= 100 * 1e4 * U.EPS64 * np.linalg.norm(Xo)
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    # the tile conversions on caller descriptors, bit-exact
    As = np.zeros((n, n), dtype=np.float32, order="F")
    A2 = np.zeros((n, n), order="F")
    s2, r2 = _seq(lib)
    S2, R2 = ctypes.byref(s2), ctypes.byref(r2)
    lib.plasma_omp_dlag2s(dA, dAs, S2, R2)
    lib.plasma_omp_sdesc2ge(dAs, As.ctypes.data, n, S2, R2)
    lib.plasma_omp_slag2d(dAs, dA, S2, R2)
    lib.plasma_omp_ddesc2ge(dA, A2.ctypes.data, n, S2, R2)
    assert lib.plasma_sequence_sync(S2) == 0
    assert np.array_equal(As, A.astype(np.float32))
    assert np.array_equal(A2, A.astype(np.float32).astype(np.float64))
    for d in (dA, dB, dX, dR, dAs, dXs):
        lib.plasma_desc_destroy(ctypes.byref(d))


def test_failure_survives_status_word_reuse(plasma, nb_guard):
    """A failing factorization's sequence, left unsynchronized while 300
    other sequences take status words (more than the 256 the library
    holds): its info must still come back from plasma_sequence_sync."""
    lib = pa.load()
    n, nb = 256, 64
    A = U.ggt_spd(n, seed=8)
    A[:, 100] = 0.0
    A[100, :] = 0.0
    dA = _general(lib, pa.PlasmaRealDouble, nb, n, n)
    bad_s, bad_r = _seq(lib)
    lib.plasma_omp_dge2desc(A.ctypes.data, n, dA, ctypes.byref(bad_s),
                            ctypes.byref(bad_r))
    lib.plasma_omp_dpotrf(L, dA, ctypes.byref(bad_s), ctypes.byref(bad_r))
    small = _general(lib, pa.PlasmaRealDouble, 32, 32, 32)
    E = np.asfortranarray(np.eye(32))
    others = [_seq(lib) for _ in range(300)]
    for s, r in others:
        lib.plasma_omp_dge2desc(E.ctypes.data, 32, small, ctypes.byref(s),
                                ctypes.byref(r))
    assert lib.plasma_sequence_sync(ctypes.byref(bad_s)) == 101
    assert bad_r.status == 101
    for s, r in others:
        assert lib.plasma_sequence_sync(ctypes.byref(s)) == 0
    lib.plasma_desc_destroy(ctypes.byref(dA))
    lib.plasma_desc_destroy(ctypes.byref(small))
