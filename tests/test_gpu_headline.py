"""GPU parity tests at the benchmark's schedules and sizes, through the C ABI.

The benchmark runs the library's autotuned schedules (PlasmaTuning, as in the
reference) replayed from captured hipGraphs (PlasmaGraphCapture), with the
FP32 trailing updates on 16-bit MFMA.  These tests run exactly that
configuration, at sizes the oracle finishes in seconds:
  * plasma_dsposv at n = 8192 and 16384, nrhs = 64, on the benchmark's
    Householder matrices (cond_2 = kappa exactly), twice in a row (capture,
    then replay): the iteration count against the oracle's (same algorithm
    and stopping rule, reference compute/zcposv.c:442-538), the solution
    against the oracle's, and the tester's scaled residual
    (test/test_zcposv.c:327-373);
  * the FP32 factor at nb = 256, update depth 4, lookahead, in graph mode,
    replayed with eager work in between, for both operand splits, against
    the oracle's FP32 tile factor (reference compute/pzpotrf.c);
  * plain FP64 plasma_dposv at n = 16384 (tuned schedule): factor against
    the oracle within 1e-12 (test/test_zpotrf.c:124-142) and the residual;
  * one 4096-system batched call at the benchmark's workload, 64 systems
    sampled against LAPACK.
"""
import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L = pa.PlasmaLower


@pytest.fixture
def tuned_graphs(plasma, nb_guard):
    pa.set_param(pa.PlasmaTuning, pa.PlasmaEnabled)
    pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaEnabled)
    yield
    pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaDisabled)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


@pytest.mark.parametrize("n,kappa", [(8192, 1e2), (8192, 1e6),
                                     (16384, 1e6)])
def test_dsposv_headline_schedule(tuned_graphs, n, kappa):
    import torch
    nrhs = 64
    A = U.householder_spd(n, kappa, seed=11)
    B = U.rhs(n, nrhs, seed=12)
    ora.set_blas(ora.MKL)
    Xo, ito, infoo = ora.dsposv(ora.LOWER, A, B, 256)
    assert infoo == 0 and 0 <= ito <= 30
    dA, dB = _dev(A), _dev(B)
    for rep in range(2):          # first call captures, second replays
        dX = torch.zeros_like(dB)
        torch.cuda.synchronize()
        info, it = pa.dsposv(L, n, nrhs, dA, n, dB, n, dX, n)
        assert info == 0
        X = dX.cpu().numpy().T
        assert abs(it - ito) <= 1, (rep, it, ito)
        err = np.linalg.norm(X - Xo) / np.linalg.norm(Xo)
        assert err <= 100 * kappa * U.EPS64, (rep, err)
        assert U.scaled_residual(A, X, B) < 10 * U.EPS64, rep


@pytest.mark.parametrize("split", ["fp16", "bf16"])
def test_spotrf_headline_schedule_graph(plasma, nb_guard, split):
    import torch
    n, nb = 4096, 256
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaUpdateDepth, 4)
    pa.set_param(pa.PlasmaLookahead, 1)
    pa.set_param(pa.PlasmaSplitUpdate, pa.PlasmaSplitFp16 if split == "fp16"
                 else pa.PlasmaSplitBf16)
    pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaEnabled)
    ora.set_blas(ora.MKL)
    try:
        for rep in range(3):
            A = U.ggt_spd(n, seed=70 + rep, dtype=np.float32)
            ref, rinfo = ora.potrf(ora.LOWER, A, nb, dtype=np.float32)
            assert rinfo == 0
            da = _dev(A)
            torch.cuda.synchronize()
            assert pa.potrf(L, n, da, n, prec="s") == 0
            F = da.cpu().numpy().T
            assert U.factor_error(F, ref, True) <= 1e-5, rep
            # eager work between replays
            pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaDisabled)
            b = U.tester_spd(300, 300)
            assert pa.potrf(L, 300, b, 300) == 0
            pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaEnabled)
    finally:
        pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaDisabled)


def test_dposv_16384_tuned(tuned_graphs):
    import torch
    n, nrhs = 16384, 64
    A = U.ggt_spd(n, seed=5)
    B = U.rhs(n, nrhs, seed=6)
    dA, dB = _dev(A), _dev(B)
    torch.cuda.synchronize()
    assert pa.posv(L, n, nrhs, dA, n, dB, n) == 0
    F = dA.cpu().numpy().T
    X = dB.cpu().numpy().T
    rc, nb = pa.get_param(pa.PlasmaNb)
    ora.set_blas(ora.MKL)
    ref, rinfo = ora.potrf(ora.LOWER, A, nb)
    assert rinfo == 0
    assert U.factor_error(np.tril(F), np.tril(ref), True) <= 1e-12
    assert U.scaled_residual(A, X, B) < 10 * U.EPS64


def test_batched_4096_systems(plasma):
    import scipy.linalg as sl
    from tests.test_gpu_batched import rbf, run
    rng = np.random.default_rng(2024)
    count = 4096
    ns = rng.integers(32, 513, count)
    mats = [rbf(int(k), rng) for k in ns]
    rhs = [rng.standard_normal((int(k), 1)) for k in ns]
    sample = rng.choice(count, 64, replace=False)
    # FP64: factor against LAPACK within 1e-12 and the FP64 residual;
    # FP32: the residual against the FP32-rounded system (the factor of an
    # RBF matrix with a 1e-3 shift differs from FP64's by about cond * u32)
    rc, info, outA, outB = run("d", L, mats, rhs, 1)
    assert rc == 0 and np.all(info == 0)
    for i in sample:
        M, b = mats[i], rhs[i]
        Lref = sl.cholesky(M, lower=True)
        Lg = np.tril(outA[i])
        assert np.linalg.norm(Lg - Lref) <= 1e-12 * np.linalg.norm(Lref), i
        assert U.scaled_residual(M, outB[i], b) < 10 * U.EPS64, i
    mats32 = [M.astype(np.float32) for M in mats]
    rhs32 = [b.astype(np.float32) for b in rhs]
    rc, info, outA, outB = run("s", L, mats32, rhs32, 1)
    assert rc == 0 and np.all(info == 0)
    for i in sample:
        res = U.scaled_residual(mats32[i].astype(np.float64),
                                outB[i].astype(np.float64),
                                rhs32[i].astype(np.float64))
        assert res < 10 * U.EPS32, i
