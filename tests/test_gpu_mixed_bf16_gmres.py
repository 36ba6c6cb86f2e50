"""GPU tests of plasma_dhposv (no reference counterpart): the FP64 solution
of A X = B from a half-precision-input factor (fp16 by default, bf16 with
PlasmaHalfInput) of the equilibrated matrix, refined by GMRES-IR, through
the C ABI.

Pinned by what fixes the answer independently: the tester's scaled
residual ||B - A X||_inf / (n ||A||_inf ||X||_inf eps64) < 10
(test/test_zcposv.c:327-360) and the FP64 drop-in plasma_dposv (a separate
code path, parity-tested against the oracle elsewhere), on matrices made
badly scaled on purpose (D C D with D log-uniform over four decades) so that
the symmetric equilibration matters.  Also: the FP64 fallback when the bf16 factor breaks
down (iter = -1, same accuracy), device-pointer arguments, unchanged
inputs, and the argument codes."""
import ctypes

import numpy as np
import pytest

import plasma_amd as pa
from tests import _util as U
from tests.test_oracle import _cond_spd

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper


def _scaled_spd(n, kappa, seed):
    """D C D with C = Q diag(s) Q^T of condition kappa, D log-uniform in
    [1e-2, 1e2]."""
    C = _cond_spd(n, kappa, seed=seed)
    d = 10.0 ** np.random.default_rng(seed + 100).uniform(-2, 2, n)
    return np.asfortranarray(C * d[:, None] * d[None, :])


def _dhposv(uplo, A, B):
    n, nrhs = B.shape
    a, b = A.copy(order="F"), B.copy(order="F")
    x = np.zeros_like(b, order="F")
    info, it = pa.dhposv(uplo, n, nrhs, a, n, b, n, x, n)
    assert np.array_equal(a, A) and np.array_equal(b, B)
    return x, it, info


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("n,nb", [(512, 128), (700, 128)])
@pytest.mark.parametrize("half", ["fp16", "bf16"])
def test_dhposv_accuracy(plasma, nb_guard, uplo, kappa, n, nb, half):
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaHalfInput, pa.PlasmaSplitFp16 if half == "fp16"
                 else pa.PlasmaSplitBf16)
    A = _scaled_spd(n, kappa, seed=3)
    B = U.rhs(n, 6, seed=4)
    X, it, info = _dhposv(uplo, A, B)
    assert info == 0
    assert it != 0
    if kappa <= 1e2:
        assert it > 0, it               # GMRES-IR, no FP64 fallback
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64
    a, xref = A.copy(order="F"), B.copy(order="F")
    assert pa.posv(uplo, n, B.shape[1], a, n, xref, n) == 0
    err = np.linalg.norm(X - xref) / np.linalg.norm(xref)
    # both solutions are backward stable, so they differ by at most about
    # cond(A) eps64; cond(A) <= kappa * 1e8 by construction (D spans four
    # decades)
    assert err < max(1e-10, kappa * 1e8 * 1e-15), err


@pytest.mark.parametrize("half", ["fp16", "bf16"])
def test_dhposv_breakdown_falls_back(plasma, nb_guard, half):
    """kappa = 1e10: the half-precision-input factor meets a pivot that is
    not positive (or refinement stalls); the FP64 factorization solves and
    iter is -1."""
    n = 384
    pa.set_param(pa.PlasmaNb, 128)
    pa.set_param(pa.PlasmaHalfInput, pa.PlasmaSplitFp16 if half == "fp16"
                 else pa.PlasmaSplitBf16)
    A = _cond_spd(n, 1e10, seed=5)
    B = U.rhs(n, 3, seed=6)
    X, it, info = _dhposv(L, A, B)
    assert info == 0
    assert it == -1
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64


def test_dhposv_device_pointers(plasma, nb_guard):
    import torch
    n, nrhs = 600, 8
    pa.set_param(pa.PlasmaNb, 128)
    A = _scaled_spd(n, 1e3, seed=9)
    B = U.rhs(n, nrhs, seed=10)
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
    dX = torch.zeros_like(dB)
    torch.cuda.synchronize()
    info, it = pa.dhposv(L, n, nrhs, dA, n, dB, n, dX, n)
    assert info == 0 and it > 0
    X = dX.cpu().numpy().T
    assert np.array_equal(dA.cpu().numpy().T, A)
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64


def test_dhposv_argument_codes(plasma):
    lib = pa.load()
    a = np.eye(4, order="F")
    b = np.ones((4, 1), order="F")
    x = np.zeros((4, 1), order="F")
    it = ctypes.c_int(0)
    p = lambda v: v.ctypes.data
    assert lib.plasma_dhposv(0, 4, 1, p(a), 4, p(b), 4, p(x), 4,
                             ctypes.byref(it)) == -1
    assert lib.plasma_dhposv(L, -1, 1, p(a), 4, p(b), 4, p(x), 4,
                             ctypes.byref(it)) == -2
    assert lib.plasma_dhposv(L, 4, -1, p(a), 4, p(b), 4, p(x), 4,
                             ctypes.byref(it)) == -3
    assert lib.plasma_dhposv(L, 4, 1, p(a), 3, p(b), 4, p(x), 4,
                             ctypes.byref(it)) == -5
    assert lib.plasma_dhposv(L, 4, 1, p(a), 4, p(b), 3, p(x), 4,
                             ctypes.byref(it)) == -7
    assert lib.plasma_dhposv(L, 4, 1, p(a), 4, p(b), 4, p(x), 3,
                             ctypes.byref(it)) == -9
    assert lib.plasma_dhposv(L, 4, 1, p(a), 4, p(b), 4, p(x), 4,
                             None) == -10
    assert lib.plasma_dhposv(L, 0, 1, p(a), 4, p(b), 4, p(x), 4,
                             ctypes.byref(it)) == 0 and it.value == 0
