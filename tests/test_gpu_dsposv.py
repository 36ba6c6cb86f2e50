"""GPU parity tests of the mixed-precision solver and the FP64 <-> FP32
conversions through the C ABI, against the CPU oracle (a restatement of
compute/zcposv.c, pinned to MKL's LAPACKE_dsposv in test_oracle.py).

Checks: the iteration count of plasma_dsposv against the oracle's (same
algorithm and stopping rule; the FP32 factors differ only in rounding, so
the counts agree within a small margin), the tester's scaled residual
(test/test_zcposv.c:327-360) below 10 eps64, the -31 FP64 fallback, the
reference's failure behaviour (FP32 breakdown -> positive info, dlag2s
overflow -> info 1, pX untouched), argument codes, and bit-exact IEEE
conversions for plasma_dlag2s / plasma_slag2d.
"""
import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U
from tests.test_oracle import _cond_spd

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper


def _dsposv(uplo, A, B, device=False):
    n, nrhs = B.shape
    if device:
        import torch
        dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        dB = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
        dX = torch.zeros_like(dB)
        torch.cuda.synchronize()
        info, it = pa.dsposv(uplo, n, nrhs, dA, n, dB, n, dX, n)
        assert np.array_equal(dA.cpu().numpy().T, A)   # pA not modified
        assert np.array_equal(dB.cpu().numpy().T, B)   # pB not modified
        return dX.cpu().numpy().T, it, info
    a, b = A.copy(order="F"), B.copy(order="F")
    x = np.zeros_like(b, order="F")
    info, it = pa.dsposv(uplo, n, nrhs, a, n, b, n, x, n)
    assert np.array_equal(a, A) and np.array_equal(b, B)
    return x, it, info


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("kappa,margin", [(1e2, 1), (1e4, 1), (1e6, 1),
                                          (3e7, None)])
@pytest.mark.parametrize("device", [False, True])
def test_dsposv_iterations_and_residual(plasma, nb_guard, uplo, kappa, margin,
                                        device):
    n, nb, nrhs = 512, 128, 5
    pa.set_param(pa.PlasmaNb, nb)
    A = _cond_spd(n, kappa, seed=7)
    B = U.rhs(n, nrhs, seed=8)
    X, it, info = _dsposv(uplo, A, B, device)
    assert info == 0
    ora.set_blas(ora.MKL)
    Xo, ito, infoo = ora.dsposv(ora.LOWER if uplo == L else ora.UPPER, A, B,
                                nb)
    assert infoo == 0
    # near kappa * u32 ~ 1 (kappa = 3e7) the count is chaotic in the
    # rounding of the FP32 factor: only convergence is compared there
    assert 0 <= it <= 30 and 0 <= ito <= 30, (it, ito)
    if margin is not None:
        assert abs(it - ito) <= margin, (it, ito)
        # both FP64-accurate: they agree to cond * eps
        err = np.linalg.norm(X - Xo) / np.linalg.norm(Xo)
        assert err <= 100 * kappa * U.EPS64, err
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb,nrhs", [(2500, 128, 37), (1900, 64, 3)])
def test_dsposv_superblock_sweeps(plasma, nb_guard, uplo, n, nb, nrhs):
    """The refinement solves run by super-blocks of 1024 rows (inverses of
    the diagonal super-blocks built once per factor): several super-blocks,
    the last one partial and ending in a short tile, more right-hand sides
    than one 32-column block.  Iterations, solution and residual against
    the oracle's tile-by-tile algorithm."""
    pa.set_param(pa.PlasmaNb, nb)
    A = _cond_spd(n, 1e4, seed=n)
    B = U.rhs(n, nrhs, seed=n + 1)
    X, it, info = _dsposv(uplo, A, B, device=True)
    assert info == 0
    ora.set_blas(ora.MKL)
    Xo, ito, infoo = ora.dsposv(ora.LOWER if uplo == L else ora.UPPER, A, B,
                                nb)
    assert infoo == 0 and 0 <= ito <= 30
    assert abs(it - ito) <= 1, (it, ito)
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) <= 100 * 1e4 * U.EPS64
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64


def test_dsposv_fp64_fallback(plasma, nb_guard):
    """kappa = 1e8: refinement with an FP32 factor stalls; after 30 steps
    iter = -31 and the FP64 factorization solves (zcposv.c:524-538)."""
    n, nb = 256, 64
    pa.set_param(pa.PlasmaNb, nb)
    A = _cond_spd(n, 1e8, seed=0)
    B = U.rhs(n, 4, seed=0)
    X, it, info = _dsposv(L, A, B)
    ora.set_blas(ora.MKL)
    _, ito, infoo = ora.dsposv(ora.LOWER, A, B, nb)
    assert info == infoo == 0
    assert it == ito == -31
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64


def test_dsposv_fp32_breakdown_fails_like_reference(plasma, nb_guard):
    """kappa = 1e9: the FP32 factor is not positive definite; the reference
    fails the sequence with that info and leaves pX untouched."""
    n, nb = 256, 64
    pa.set_param(pa.PlasmaNb, nb)
    A = _cond_spd(n, 1e9, seed=0)
    B = U.rhs(n, 2, seed=0)
    a, b = A.copy(order="F"), B.copy(order="F")
    x = np.full_like(b, 7.0, order="F")
    info, it = pa.dsposv(L, n, 2, a, n, b, n, x, n)
    _, _, infoo = ora.dsposv(ora.LOWER, A, B, nb)
    assert info > 0 and infoo > 0   # the failing pivot depends on rounding
    assert np.all(x == 7.0)


def test_dsposv_overflow_fails_with_info_1(plasma, nb_guard):
    n = 64
    pa.set_param(pa.PlasmaNb, 32)
    A = U.ggt_spd(n)
    B = U.rhs(n, 2)
    B[3, 1] = 1e39
    x = np.zeros_like(B, order="F")
    info, it = pa.dsposv(L, n, 2, A.copy(order="F"), n, B.copy(order="F"), n,
                         x, n)
    assert info == 1
    assert np.all(x == 0.0)


def test_dsposv_argument_codes(plasma):
    a = np.eye(4, order="F")
    b = np.ones((4, 2), order="F")
    x = np.zeros((4, 2), order="F")
    assert pa.dsposv(999, 4, 2, a, 4, b, 4, x, 4)[0] == -1
    assert pa.dsposv(L, -1, 2, a, 4, b, 4, x, 4)[0] == -2
    assert pa.dsposv(L, 4, -1, a, 4, b, 4, x, 4)[0] == -3
    assert pa.dsposv(L, 4, 2, a, 3, b, 4, x, 4)[0] == -5
    assert pa.dsposv(L, 4, 2, a, 4, b, 3, x, 4)[0] == -7
    assert pa.dsposv(L, 4, 2, a, 4, b, 4, x, 3)[0] == -9
    assert pa.dsposv(L, 4, 0, a, 4, b, 4, x, 4) == (0, 0)


@pytest.mark.parametrize("m,n", [(1, 1), (300, 70), (513, 257)])
def test_lag2_bit_exact(plasma, m, n):
    lib = pa.load()
    rng = np.random.default_rng(m + n)
    A = np.asfortranarray(rng.standard_normal((m + 2, n)) * 1e5)
    S = np.zeros((m + 1, n), dtype=np.float32, order="F")
    assert lib.plasma_dlag2s(m, n, A.ctypes.data, m + 2, S.ctypes.data,
                             m + 1) == 0
    assert np.array_equal(S[:m], A[:m].astype(np.float32))
    D = np.zeros((m + 3, n), order="F")
    assert lib.plasma_slag2d(m, n, S.ctypes.data, m + 1, D.ctypes.data,
                             m + 3) == 0
    assert np.array_equal(D[:m], S[:m].astype(np.float64))
    A[m // 2, n // 2] = -np.inf
    assert lib.plasma_dlag2s(m, n, A.ctypes.data, m + 2, S.ctypes.data,
                             m + 1) == 1
    assert lib.plasma_dlag2s(-1, n, A.ctypes.data, m + 2, S.ctypes.data,
                             m + 1) == -1
    assert lib.plasma_dlag2s(m, n, A.ctypes.data, 0, S.ctypes.data,
                             m + 1) == -4
    assert lib.plasma_slag2d(m, n, S.ctypes.data, 0, D.ctypes.data,
                             m + 3) == -4
