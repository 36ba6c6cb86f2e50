"""GPU parity tests of the public BLAS-3 front ends through the C ABI:
plasma_{d,s}gemm, plasma_dsyrk and plasma_{d,s}trsm (and their tile forms)
against the oracle's restatements of compute/pzgemm.c, pzsyrk.c and
pztrsm.c (pinned to SciPy's BLAS in test_oracle_blas3.py), at the same tile
size, on host and device arrays, with edge tiles (sizes not multiples of
nb), every transposition / side / triangle / diagonal kind, and the
reference's argument codes.

Tolerances: the device sums each K range in a different order (MFMA blocks)
than the oracle's tile loop, so results agree to rounding:
|C - C_oracle| <= 4 k eps (|alpha| |op(A)| |op(B)| + |beta| |C|) elementwise
for gemm / syrk; trsm (by tile inverses on the device, substitution in the
oracle) within 1e-11 relative on well-conditioned triangles (FP64) and
1e-4 (FP32)."""
import ctypes
import itertools

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

NT, T = pa.PlasmaNoTrans, pa.PlasmaTrans
OMAP = {pa.PlasmaNoTrans: ora.NOTRANS, pa.PlasmaTrans: ora.TRANS,
        pa.PlasmaConjTrans: ora.TRANS, pa.PlasmaLower: ora.LOWER,
        pa.PlasmaUpper: ora.UPPER, pa.PlasmaLeft: ora.LEFT,
        pa.PlasmaRight: ora.RIGHT, pa.PlasmaUnit: ora.UNIT,
        pa.PlasmaNonUnit: ora.NONUNIT}


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def _host(t):
    return np.asfortranarray(t.cpu().numpy().T)


def _gemm_bound(alpha, opa, opb, beta, C, eps):
    k = opa.shape[1]
    return 4 * max(k, 1) * eps * (abs(alpha) * np.abs(opa) @ np.abs(opb)
                                  + abs(beta) * np.abs(C)) + 1e-300


@pytest.mark.parametrize("ta,tb", list(itertools.product(
    [NT, T, pa.PlasmaConjTrans], [NT, T])))
@pytest.mark.parametrize("device", [False, True])
def test_dgemm(plasma, nb_guard, ta, tb, device):
    import torch
    nb = 128
    pa.set_param(pa.PlasmaNb, nb)
    rng = np.random.default_rng(ta * 3 + tb)
    m, n, k = 300, 270, 333
    A = np.asfortranarray(rng.standard_normal((m, k) if ta == NT else (k, m)))
    B = np.asfortranarray(rng.standard_normal((k, n) if tb == NT else (n, k)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    alpha, beta = 1.5, -0.5
    lib = pa.load()
    if device:
        dA, dB, dC = _dev(A), _dev(B), _dev(C)
        torch.cuda.synchronize()
        rc = lib.plasma_dgemm(ta, tb, m, n, k, alpha, dA.data_ptr(),
                              A.shape[0], dB.data_ptr(), B.shape[0], beta,
                              dC.data_ptr(), m)
        got = _host(dC)
    else:
        got = C.copy(order="F")
        rc = lib.plasma_dgemm(ta, tb, m, n, k, alpha, A.ctypes.data,
                              A.shape[0], B.ctypes.data, B.shape[0], beta,
                              got.ctypes.data, m)
    assert rc == 0
    ora.set_blas(ora.MKL)
    ref = ora.gemm(OMAP[ta], OMAP[tb], alpha, A, B, beta, C, nb)
    opa = A if ta == NT else A.T
    opb = B if tb == NT else B.T
    assert np.all(np.abs(got - ref) <= _gemm_bound(alpha, opa, opb, beta, C,
                                                  U.EPS64))


def test_dgemm_scalars_and_codes(plasma, nb_guard):
    pa.set_param(pa.PlasmaNb, 64)
    lib = pa.load()
    rng = np.random.default_rng(1)
    m, n, k = 100, 90, 70
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((k, n)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    # beta = 0: C is not read (NaN input gives a finite result)
    got = np.full((m, n), np.nan, order="F")
    assert lib.plasma_dgemm(NT, NT, m, n, k, 1.0, A.ctypes.data, m,
                            B.ctypes.data, k, 0.0, got.ctypes.data, m) == 0
    assert np.abs(got - A @ B).max() <= 1e-13 * np.abs(A @ B).max() * k
    # alpha = 0: C = beta C, operands not read
    An = np.full((m, k), np.nan, order="F")
    got = C.copy(order="F")
    assert lib.plasma_dgemm(NT, NT, m, n, k, 0.0, An.ctypes.data, m,
                            B.ctypes.data, k, 2.0, got.ctypes.data, m) == 0
    assert np.array_equal(got, 2.0 * C)
    # argument codes (reference compute/zgemm.c:120-170)
    p = C.ctypes.data
    assert lib.plasma_dgemm(7, NT, m, n, k, 1.0, p, m, p, k, 1.0, p, m) == -1
    assert lib.plasma_dgemm(NT, 7, m, n, k, 1.0, p, m, p, k, 1.0, p, m) == -2
    assert lib.plasma_dgemm(NT, NT, -1, n, k, 1.0, p, m, p, k, 1.0, p, m) == -3
    assert lib.plasma_dgemm(NT, NT, m, -1, k, 1.0, p, m, p, k, 1.0, p, m) == -4
    assert lib.plasma_dgemm(NT, NT, m, n, -1, 1.0, p, m, p, k, 1.0, p, m) == -5
    assert lib.plasma_dgemm(NT, NT, m, n, k, 1.0, p, m - 1, p, k, 1.0, p,
                            m) == -8
    assert lib.plasma_dgemm(NT, NT, m, n, k, 1.0, p, m, p, k - 1, 1.0, p,
                            m) == -10
    assert lib.plasma_dgemm(NT, NT, m, n, k, 1.0, p, m, p, k, 1.0, p,
                            m - 1) == -13
    # quick return: nothing touched
    assert lib.plasma_dgemm(NT, NT, 0, n, k, 1.0, p, 1, p, k, 1.0, p, 1) == 0


def test_sgemm(plasma, nb_guard):
    pa.set_param(pa.PlasmaNb, 96)
    lib = pa.load()
    rng = np.random.default_rng(4)
    m, n, k = 257, 130, 200
    A = np.asfortranarray(rng.standard_normal((k, m)), dtype=np.float32)
    B = np.asfortranarray(rng.standard_normal((n, k)), dtype=np.float32)
    C = np.asfortranarray(rng.standard_normal((m, n)), dtype=np.float32)
    got = C.copy(order="F")
    assert lib.plasma_sgemm(T, T, m, n, k, 0.5, A.ctypes.data, k,
                            B.ctypes.data, n, 2.0, got.ctypes.data, m) == 0
    ora.set_blas(ora.MKL)
    ref = ora.gemm(ora.TRANS, ora.TRANS, 0.5, A, B, 2.0, C, 96,
                   dtype=np.float32)
    bound = _gemm_bound(0.5, A.T.astype(np.float64), B.T.astype(np.float64),
                        2.0, C.astype(np.float64), U.EPS32)
    assert np.all(np.abs(got.astype(np.float64) - ref) <= 2 * bound)


@pytest.mark.parametrize("uplo", [pa.PlasmaLower, pa.PlasmaUpper])
@pytest.mark.parametrize("trans", [NT, T])
def test_dsyrk(plasma, nb_guard, uplo, trans):
    nb = 128
    pa.set_param(pa.PlasmaNb, nb)
    lib = pa.load()
    rng = np.random.default_rng(uplo + trans)
    n, k = 333, 290
    A = np.asfortranarray(rng.standard_normal((n, k) if trans == NT
                                              else (k, n)))
    C = np.asfortranarray(rng.standard_normal((n, n)))
    got = C.copy(order="F")
    assert lib.plasma_dsyrk(uplo, trans, n, k, 0.75, A.ctypes.data,
                            A.shape[0], 0.5, got.ctypes.data, n) == 0
    ora.set_blas(ora.MKL)
    ref = ora.syrk(OMAP[uplo], OMAP[trans], 0.75, A, 0.5, C, nb)
    opa = A if trans == NT else A.T
    bound = _gemm_bound(0.75, opa, opa.T, 0.5, C, U.EPS64)
    tri = np.tril if uplo == pa.PlasmaLower else np.triu
    # the stored triangle within the bound (tri of a difference, not of a
    # boolean mask, whose other triangle would read False)
    assert np.all(tri(np.abs(got - ref) - bound) <= 0)
    # the other triangle (strictly) is untouched, bit for bit
    other = np.triu if uplo == pa.PlasmaLower else np.tril
    off = 1 if uplo == pa.PlasmaLower else -1
    assert np.array_equal(other(got, off), other(C, off))
    # argument codes (reference compute/zsyrk.c:102-138)
    p = C.ctypes.data
    assert lib.plasma_dsyrk(7, trans, n, k, 1.0, p, n, 1.0, p, n) == -1
    assert lib.plasma_dsyrk(uplo, pa.PlasmaConjTrans, n, k, 1.0, p, n, 1.0,
                            p, n) == -2
    assert lib.plasma_dsyrk(uplo, trans, -1, k, 1.0, p, n, 1.0, p, n) == -3
    assert lib.plasma_dsyrk(uplo, trans, n, -1, 1.0, p, n, 1.0, p, n) == -4
    assert lib.plasma_dsyrk(uplo, trans, n, k, 1.0, p, 0, 1.0, p, n) == -7
    assert lib.plasma_dsyrk(uplo, trans, n, k, 1.0, p, n + k, 1.0, p,
                            n - 1) == -10


def _tri(n, lower, rng, dtype=np.float64):
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    a = np.tril(a) if lower else np.triu(a)
    a[np.diag_indices(n)] = 2.0 + rng.random(n)
    # garbage in the other triangle: must not be read
    a += (np.triu if lower else np.tril)(rng.standard_normal((n, n)),
                                         1 if lower else -1) * 100
    return np.asfortranarray(a, dtype=dtype)


CASES = list(itertools.product([pa.PlasmaLeft, pa.PlasmaRight],
                               [pa.PlasmaLower, pa.PlasmaUpper],
                               [NT, T], [pa.PlasmaNonUnit, pa.PlasmaUnit]))


@pytest.mark.parametrize("side,uplo,trans,diag", CASES)
def test_dtrsm(plasma, nb_guard, side, uplo, trans, diag):
    import torch
    nb = 128
    pa.set_param(pa.PlasmaNb, nb)
    lib = pa.load()
    rng = np.random.default_rng(side + uplo + trans + diag)
    m, n = 400, 270
    na = m if side == pa.PlasmaLeft else n
    A = _tri(na, uplo == pa.PlasmaLower, rng)
    B = np.asfortranarray(rng.standard_normal((m, n)))
    dA, dB = _dev(A), _dev(B)
    torch.cuda.synchronize()
    assert lib.plasma_dtrsm(side, uplo, trans, diag, m, n, -1.5,
                            dA.data_ptr(), na, dB.data_ptr(), m) == 0
    got = _host(dB)
    ora.set_blas(ora.MKL)
    ref = ora.trsm(OMAP[side], OMAP[uplo], OMAP[trans], OMAP[diag], -1.5, A, B,
                   nb)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
    assert np.array_equal(_host(dA), A)          # A is not modified


def test_trsm_codes_and_single(plasma, nb_guard):
    pa.set_param(pa.PlasmaNb, 64)
    lib = pa.load()
    rng = np.random.default_rng(2)
    m, n = 150, 70
    A = _tri(m, True, rng, np.float32)
    B = np.asfortranarray(rng.standard_normal((m, n)), dtype=np.float32)
    got = B.copy(order="F")
    assert lib.plasma_strsm(pa.PlasmaLeft, pa.PlasmaLower, T,
                            pa.PlasmaNonUnit, m, n, 1.0, A.ctypes.data, m,
                            got.ctypes.data, m) == 0
    ora.set_blas(ora.MKL)
    ref = ora.trsm(ora.LEFT, ora.LOWER, ora.TRANS, ora.NONUNIT, 1.0, A, B, 64,
                   dtype=np.float32)
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()
    p = B.ctypes.data
    L, Lw = pa.PlasmaLeft, pa.PlasmaLower
    assert lib.plasma_dtrsm(7, Lw, NT, pa.PlasmaUnit, m, n, 1.0, p, m, p,
                            m) == -1
    assert lib.plasma_dtrsm(L, 7, NT, pa.PlasmaUnit, m, n, 1.0, p, m, p,
                            m) == -2
    assert lib.plasma_dtrsm(L, Lw, 7, pa.PlasmaUnit, m, n, 1.0, p, m, p,
                            m) == -3
    assert lib.plasma_dtrsm(L, Lw, NT, 7, m, n, 1.0, p, m, p, m) == -4
    assert lib.plasma_dtrsm(L, Lw, NT, pa.PlasmaUnit, -1, n, 1.0, p, m, p,
                            m) == -5
    assert lib.plasma_dtrsm(L, Lw, NT, pa.PlasmaUnit, m, -1, 1.0, p, m, p,
                            m) == -6
    assert lib.plasma_dtrsm(L, Lw, NT, pa.PlasmaUnit, m, n, 1.0, p, m - 1, p,
                            m) == -8
    assert lib.plasma_dtrsm(L, Lw, NT, pa.PlasmaUnit, m, n, 1.0, p, m, p,
                            m - 1) == -10


def test_omp_dgemm_descriptors(plasma, nb_guard):
    """The tile form on caller descriptors, chained with a tile trsm in one
    sequence: C = A B, then C <- L^{-1} C."""
    lib = pa.load()
    nb = 128
    rng = np.random.default_rng(8)
    m, n, k = 384, 200, 256
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((k, n)))
    Lm = _tri(m, True, rng)
    out = np.zeros((m, n), order="F")

    def general(mm, nn):
        d = pa.plasma_desc_t()
        assert lib.plasma_desc_general_create(pa.PlasmaRealDouble, nb, nb, mm,
                                              nn, 0, 0, mm, nn,
                                              ctypes.byref(d)) == 0
        return d
    dA, dB, dC, dL = general(m, k), general(k, n), general(m, n), general(m, m)
    s, r = pa.plasma_sequence_t(), pa.plasma_request_t()
    lib.plasma_sequence_init(ctypes.byref(s))
    lib.plasma_request_init(ctypes.byref(r))
    S, R = ctypes.byref(s), ctypes.byref(r)
    lib.plasma_omp_dge2desc(A.ctypes.data, m, dA, S, R)
    lib.plasma_omp_dge2desc(B.ctypes.data, k, dB, S, R)
    lib.plasma_omp_dge2desc(Lm.ctypes.data, m, dL, S, R)
    lib.plasma_omp_dgemm(NT, NT, 1.0, dA, dB, 0.0, dC, S, R)
    lib.plasma_omp_dtrsm(pa.PlasmaLeft, pa.PlasmaLower, NT, pa.PlasmaNonUnit,
                         1.0, dL, dC, S, R)
    lib.plasma_omp_ddesc2ge(dC, out.ctypes.data, m, S, R)
    assert lib.plasma_sequence_sync(S) == 0
    ora.set_blas(ora.MKL)
    C = ora.gemm(ora.NOTRANS, ora.NOTRANS, 1.0, A, B, 0.0, np.zeros((m, n)), nb)
    ref = ora.trsm(ora.LEFT, ora.LOWER, ora.NOTRANS, ora.NONUNIT, 1.0, Lm, C,
                   nb)
    assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max()
    for d in (dA, dB, dC, dL):
        lib.plasma_desc_destroy(ctypes.byref(d))
