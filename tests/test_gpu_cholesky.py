"""GPU parity tests of the FP64/FP32 Cholesky path through the C ABI.

The checks are the reference tester's (test/test_zpotrf.c:124-152,
test_zposv.c:144-168, test_zpotrs.c:308-331) against the CPU oracle
(a restatement of PLASMA's tile algorithm, itself pinned to LAPACK in
test_oracle.py):
  * potrf: ||L - L_oracle||_F / ||L_oracle||_F <= 1e-12, and the entries the
    factorization must not touch are bit-identical to the input;
  * posv / potrs: scaled residual ||B - AX||_inf / (n ||A||_inf ||X||_inf)
    < 10 * eps64 (stricter than the tester's 50 eps);
  * a non positive definite input returns LAPACK's info and leaves pA as
    it was (the reference skips its desc2tr copy on a failed sequence).
"""
import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper

CASES = [(1, 16), (15, 16), (100, 16), (256, 256), (257, 64), (300, 128),
         (1000, 256), (1536, 512), (2048, 256)]


def _oracle_factor(uplo, a, nb):
    ora.set_blas(ora.MKL)
    f, info = ora.potrf(uplo, a, nb)
    return f, info


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb", CASES)
def test_dpotrf_host(plasma, nb_guard, uplo, n, nb):
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    lda = n + 3
    A = U.tester_spd(n, lda)
    a = A.copy(order="F")
    info = pa.potrf(uplo, n, a, lda)
    assert info == 0
    ref, rinfo = _oracle_factor(uplo, A[:n, :n], nb)
    assert rinfo == 0
    err = U.factor_error(a[:n, :n], ref, uplo == L)
    assert err <= 1e-12, err
    # padding rows untouched
    assert np.array_equal(a[n:, :], A[n:, :])


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb", [(100, 16), (1000, 256), (2048, 256)])
def test_dpotrf_device(plasma, nb_guard, uplo, n, nb):
    import torch
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=n)
    # column-major on the device: tensor row j holds column j
    d = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    torch.cuda.synchronize()
    info = pa.potrf(uplo, n, d, n)
    assert info == 0
    got = d.cpu().numpy().T
    ref, _ = _oracle_factor(uplo, A, nb)
    assert U.factor_error(got, ref, uplo == L) <= 1e-12


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb,nrhs", [(15, 16, 3), (300, 128, 17),
                                       (1000, 256, 64), (2048, 256, 64),
                                       (1536, 512, 100)])
def test_dposv(plasma, nb_guard, uplo, n, nb, nrhs):
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=2 * n + 1)
    B = U.rhs(n, nrhs, seed=n)
    a, b = A.copy(order="F"), B.copy(order="F")
    assert pa.posv(uplo, n, nrhs, a, n, b, n) == 0
    res = U.scaled_residual(A, b, B)
    assert res < 10 * U.EPS64, res / U.EPS64
    _, xref, _ = ora.posv(uplo, A, B, nb)
    assert np.abs(b - xref).max() <= 1e-10 * np.abs(xref).max()


@pytest.mark.parametrize("uplo", [L, Up])
def test_dpotrs(plasma, nb_guard, uplo):
    n, nb, nrhs = 777, 128, 9
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    A = U.tester_spd(n)
    f = A.copy(order="F")
    assert pa.potrf(uplo, n, f, n) == 0
    B = U.rhs(n, nrhs, seed=5)
    b = B.copy(order="F")
    assert pa.potrs(uplo, n, nrhs, f, n, b, n) == 0
    assert U.scaled_residual(A, b, B) < 10 * U.EPS64


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("zerocol", [0, 37, 255, 256, 299])
def test_dpotrf_not_pd(plasma, nb_guard, uplo, zerocol):
    """zerocol (test/test_zpotrf.c:97-99): info = LAPACK's, pA unchanged."""
    import scipy.linalg as sl
    n, nb = 300, 128
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    A = U.tester_spd(n)
    A[:, zerocol] = 0.0
    a = A.copy(order="F")
    info = pa.potrf(uplo, n, a, n)
    lapinfo = sl.lapack.dpotrf(A, lower=(uplo == L))[1]
    assert info == lapinfo == zerocol + 1
    assert np.array_equal(a, A)
    _, oinfo = ora.potrf(uplo, A, nb)
    assert oinfo == info


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb,depth", [(1000, 128, 2), (2048, 256, 1),
                                        (1500, 96, 3), (700, 512, 1),
                                        (1300, 256, 4)])
@pytest.mark.parametrize("mode", ["bf16", "fp16"])
def test_spotrf_split_update(plasma, nb_guard, uplo, n, nb, depth, mode):
    """FP32 trailing updates on 16-bit MFMA from exact operand splits
    (three bf16 terms and six products, or row-scaled two-term fp16 and
    three products): the factor matches the FP32-MFMA one to FP32 accuracy
    (error vs the oracle within 4x of the plain FP32 path's, and <= 1e-5)."""
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaUpdateDepth, depth)
    A = U.ggt_spd(n, seed=n, dtype=np.float32)
    ref, _ = ora.potrf(ora.LOWER if uplo == L else ora.UPPER, A, nb,
                       dtype=np.float32)
    errs = []
    split = pa.PlasmaSplitBf16 if mode == "bf16" else pa.PlasmaSplitFp16
    for flag in (pa.PlasmaDisabled, split):
        pa.set_param(pa.PlasmaSplitUpdate, flag)
        a = A.copy(order="F")
        assert pa.potrf(uplo, n, a, n, prec="s") == 0
        errs.append(U.factor_error(a, ref, uplo == L))
    assert errs[1] <= 1e-5 and errs[1] <= 4 * errs[0] + 1e-7, errs


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("lookahead", [0, 1])
def test_spotrf_fp16_split_badly_scaled(plasma, nb_guard, uplo, lookahead):
    """Rows scaled over twelve orders of magnitude (A = D G D, D log-spaced
    in [1e-6, 1e6]): the fp16 terms are formed after scaling each row of
    the factor by a power of two that bounds it by 1, so the factor keeps
    FP32 accuracy row by row.  Checked against the oracle's FP32 factor
    row-wise relative to each row's scale, next to the FP32-MFMA path."""
    n, nb = 1024, 256
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaLookahead, lookahead)
    d = np.logspace(-6, 6, n)
    rng = np.random.default_rng(5)
    perm = rng.permutation(n)
    d = d[perm]
    A64 = U.ggt_spd(n, seed=3) * d[:, None] * d[None, :]
    A = np.asfortranarray(A64.astype(np.float32))
    ref, _ = ora.potrf(ora.LOWER if uplo == L else ora.UPPER, A, nb,
                       dtype=np.float32)
    low = uplo == L
    tri = np.tril if low else np.triu
    dscale = d if low else d            # row r of L (col r of U) ~ d_r
    errs = []
    for flag in (pa.PlasmaDisabled, pa.PlasmaSplitFp16):
        pa.set_param(pa.PlasmaSplitUpdate, flag)
        a = A.copy(order="F")
        assert pa.potrf(uplo, n, a, n, prec="s") == 0
        diff = tri(a.astype(np.float64) - ref.astype(np.float64))
        if not low:
            diff = diff.T
        # row r of L scales like d_r: relative error per row
        rel = np.abs(diff).max(axis=1) / dscale
        errs.append(rel.max())
    assert errs[1] <= 1e-4 and errs[1] <= 8 * errs[0] + 1e-7, errs


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("lookahead", [0, 1, 3])
@pytest.mark.parametrize("n,nb", [(1000, 128), (1300, 256)])
def test_spotrf_panel_schedules(plasma, nb_guard, uplo, lookahead, n, nb):
    """FP32 solves its panels by 32-column block substitution
    (diagonal-block inverses only), with and without lookahead (FP32
    clamps lookahead 3 to 2): against the oracle, and sposv's residual
    (the sweeps need the whole W_kk, formed after the factorization)."""
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaLookahead, lookahead)
    A = U.ggt_spd(n, seed=n + lookahead, dtype=np.float32)
    ref, _ = ora.potrf(ora.LOWER if uplo == L else ora.UPPER, A, nb,
                       dtype=np.float32)
    a = A.copy(order="F")
    assert pa.potrf(uplo, n, a, n, prec="s") == 0
    assert U.factor_error(a, ref, uplo == L) <= 1e-5
    B = U.rhs(n, 5, seed=4, dtype=np.float32)
    a2, b = A.copy(order="F"), B.copy(order="F")
    assert pa.posv(uplo, n, 5, a2, n, b, n, prec="s") == 0
    res = U.scaled_residual(A.astype(np.float64), b.astype(np.float64),
                            B.astype(np.float64))
    assert res < 10 * U.EPS32


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("prec", ["d", "s"])
def test_potrs_wide_rhs(plasma, nb_guard, uplo, prec):
    """Sweeps with >= 1024 right-hand sides run as 128 x 128 block GEMMs:
    edge tiles in both dimensions, both triangles, against the oracle."""
    n, nb, nrhs = 700, 128, 1100
    pa.set_param(pa.PlasmaNb, nb)
    dt = np.float64 if prec == "d" else np.float32
    A = U.ggt_spd(n, seed=11)
    F, info = ora.potrf(ora.LOWER if uplo == L else ora.UPPER, A, nb)
    assert info == 0
    B = U.rhs(n, nrhs, seed=12)
    ref, _ = ora.potrs(ora.LOWER if uplo == L else ora.UPPER, F, B, nb)
    f = np.asfortranarray(F.astype(dt))
    b = np.asfortranarray(B.astype(dt))
    assert pa.potrs(uplo, n, nrhs, f, n, b, n, prec=prec) == 0
    err = np.linalg.norm(b - ref) / np.linalg.norm(ref)
    assert err < (1e-12 if prec == "d" else 1e-4), err


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("zerocol", [0, 37, 127, 128, 200])
def test_spotrf_not_pd(plasma, nb_guard, uplo, zerocol):
    """FP32 tiles of order <= 128 are factored in LDS: LAPACK's info and
    pA unchanged, as for the global-memory form."""
    n, nb = 260, 128
    pa.set_param(pa.PlasmaNb, nb)
    A = U.tester_spd(n, dtype=np.float32)
    A[:, zerocol] = 0.0
    a = A.copy(order="F")
    assert pa.potrf(uplo, n, a, n, prec="s") == zerocol + 1
    assert np.array_equal(a, A)


def test_argument_codes(plasma):
    a = np.eye(4, order="F")
    b = np.ones((4, 2), order="F")
    assert pa.potrf(999, 4, a, 4) == -1
    assert pa.potrf(L, -1, a, 4) == -2
    assert pa.potrf(L, 4, a, 3) == -4
    assert pa.posv(999, 4, 2, a, 4, b, 4) == -1
    assert pa.posv(L, -1, 2, a, 4, b, 4) == -2
    assert pa.posv(L, 4, -1, a, 4, b, 4) == -3
    assert pa.posv(L, 4, 2, a, 3, b, 4) == -5
    assert pa.posv(L, 4, 2, a, 4, b, 3) == -7
    assert pa.potrs(L, 4, 2, a, 4, b, 3) == -7
    # quick returns
    assert pa.potrf(L, 0, a, 1) == 0
    assert pa.posv(L, 4, 0, a, 4, b, 4) == 0


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb", [(100, 16), (515, 96), (1000, 128),
                                  (1000, 256), (1536, 512)])
def test_spotrf_sposv(plasma, nb_guard, uplo, n, nb):
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=7, dtype=np.float32)
    a = A.copy(order="F")
    assert pa.potrf(uplo, n, a, n, prec="s") == 0
    ref, _ = ora.potrf(uplo, A, nb, dtype=np.float32)
    assert U.factor_error(a, ref, uplo == L) <= 1e-5
    B = U.rhs(n, 8, seed=3, dtype=np.float32)
    a2, b = A.copy(order="F"), B.copy(order="F")
    assert pa.posv(uplo, n, 8, a2, n, b, n, prec="s") == 0
    res = U.scaled_residual(A.astype(np.float64), b.astype(np.float64),
                            B.astype(np.float64))
    assert res < 10 * U.EPS32


@pytest.mark.parametrize("lookahead", [0, 1, 2, 4])
def test_dposv_8192(plasma, nb_guard, lookahead):
    """A full-size-class solve on device memory: residual and the factor
    against the oracle (MKL tiles on the host)."""
    import torch
    n, nb, nrhs = 8192, 256, 64
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaLookahead, lookahead)
    A = U.ggt_spd(n, seed=11)
    B = U.rhs(n, nrhs, seed=12)
    da = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
    torch.cuda.synchronize()
    assert pa.posv(L, n, nrhs, da, n, db, n) == 0
    X = db.cpu().numpy().T
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    F = da.cpu().numpy().T
    ref, _ = _oracle_factor(L, A, nb)
    assert U.factor_error(F, ref, True) <= 1e-12


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("depth", [1, 2, 3, 4])
@pytest.mark.parametrize("lookahead", [0, 1, 2, 3])
@pytest.mark.parametrize("n,nb", [(1000, 128), (1152, 128), (300, 64)])
def test_dpotrf_update_depth(plasma, nb_guard, uplo, depth, lookahead, n,
                             nb):
    """Trailing updates by groups of 1-4 panels (K = depth * nb), without
    lookahead and with the panel stream 1-3 groups ahead, including a short
    last tile row (n % nb != 0), tile counts that are not a multiple of the
    group size and more lookahead than groups: the factor against the
    oracle and the untouched triangle."""
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaLookahead, lookahead)
    pa.set_param(pa.PlasmaUpdateDepth, depth)
    A = U.tester_spd(n, n)
    a = A.copy(order="F")
    assert pa.potrf(uplo, n, a, n) == 0
    ref, rinfo = _oracle_factor(uplo, A, nb)
    assert rinfo == 0
    assert U.factor_error(a, ref, uplo == L) <= 1e-12
    other = np.triu(A, 1) if uplo == L else np.tril(A, -1)
    got = np.triu(a, 1) if uplo == L else np.tril(a, -1)
    assert np.array_equal(got, other)


@pytest.mark.parametrize("depth", [1, 3])
def test_dposv_8192_depth(plasma, nb_guard, depth):
    """n = 8192 with lookahead at update depths other than the default."""
    import torch
    n, nb, nrhs = 8192, 256, 64
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaLookahead, 1)
    pa.set_param(pa.PlasmaUpdateDepth, depth)
    A = U.ggt_spd(n, seed=13)
    B = U.rhs(n, nrhs, seed=14)
    da = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
    torch.cuda.synchronize()
    assert pa.posv(L, n, nrhs, da, n, db, n) == 0
    X = db.cpu().numpy().T
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    F = da.cpu().numpy().T
    ref, _ = _oracle_factor(L, A, nb)
    assert U.factor_error(F, ref, True) <= 1e-12
