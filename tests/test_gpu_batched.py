"""Batched SPD solves of uneven order (plasma_{d,s}posv_vbatched,
plasma_dpotrf_vbatched): every system against LAPACK on that matrix alone
(MKL's dpotrf through the oracle, one tile covering the whole matrix), the
tester's scaled residual
per system, LAPACK's info for a system that is not positive definite (the
others unaffected), both triangles, host and device metadata arrays, and
the argument codes.  Inputs: RBF kernel matrices exp(-|x-y|^2 / (2 0.2^2))
on points in [0,1]^3 plus 1e-3 I, the benchmark's batched workload at a
reduced count."""
import ctypes

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper


def rbf(n, rng):
    x = rng.random((n, 3))
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    return np.exp(-d2 / (2 * 0.2 ** 2)) + 1e-3 * np.eye(n)


def pack(mats, rhs, dtype):
    import torch
    aoff, boff, a, b = [], [], [], []
    pa_, pb = 0, 0
    for M, R in zip(mats, rhs):
        aoff.append(pa_)
        boff.append(pb)
        a.append(np.asfortranarray(M).ravel(order="F"))
        b.append(np.asfortranarray(R).ravel(order="F"))
        pa_ += M.size
        pb += R.size
    A = torch.from_numpy(np.concatenate(a + [np.zeros(1)]).astype(dtype)).cuda()
    B = torch.from_numpy(np.concatenate(b + [np.zeros(1)]).astype(dtype)).cuda()
    return A, B, np.array(aoff, dtype=np.uint64), np.array(boff, dtype=np.uint64)


def run(prec, uplo, mats, rhs, nrhs, device_meta=False):
    import torch
    dt = np.float64 if prec == "d" else np.float32
    A, B, aoff, boff = pack(mats, rhs, dt)
    n = np.array([M.shape[0] for M in mats], dtype=np.int32)
    info = np.full(len(mats), -7, dtype=np.int32)
    if device_meta:
        nd = torch.from_numpy(n).cuda()
        ad = torch.from_numpy(aoff.view(np.int64)).cuda()
        bd = torch.from_numpy(boff.view(np.int64)).cuda()
        idv = torch.from_numpy(info).cuda()
        args = (nd.data_ptr(), A.data_ptr(), ad.data_ptr(), nrhs, B.data_ptr(),
                bd.data_ptr(), idv.data_ptr())
    else:
        args = (n.ctypes.data, A.data_ptr(), aoff.ctypes.data, nrhs,
                B.data_ptr(), boff.ctypes.data, info.ctypes.data)
    fn = getattr(pa.load(), "plasma_%sposv_vbatched" % prec)
    rc = fn(uplo, len(mats), *args)
    torch.cuda.synchronize()
    if device_meta:
        info = idv.cpu().numpy()
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    outA = [Ah[int(o):int(o) + M.size].reshape(M.shape, order="F")
            for o, M in zip(aoff, mats)]
    outB = [Bh[int(o):int(o) + R.size].reshape(R.shape, order="F")
            for o, R in zip(boff, rhs)]
    return rc, info, outA, outB


def lapack_chol(M):
    """Lower Cholesky factor of M by LAPACK dpotrf (MKL)."""
    ora.set_blas(ora.MKL)
    F, info = ora.potrf(ora.LOWER, M, max(M.shape[0], 1))
    assert info == 0
    return np.tril(F)


def systems(count, seed, sizes=None):
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = list(rng.integers(32, 513, count))
    mats = [rbf(int(k), rng) for k in sizes]
    return mats, rng


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("device_meta", [False, True])
def test_dposv_vbatched(plasma, uplo, device_meta):
    sizes = [0, 1, 2, 31, 32, 33, 64, 100, 255, 256, 257, 511, 512] + \
        list(np.random.default_rng(3).integers(32, 513, 40))
    mats, rng = systems(0, 1, sizes)
    nrhs = 2
    rhs = [rng.standard_normal((M.shape[0], nrhs)) for M in mats]
    rc, info, outA, outB = run("d", uplo, mats, rhs, nrhs, device_meta)
    assert rc == 0
    assert np.all(info == 0)
    for M, R, F, X in zip(mats, rhs, outA, outB):
        k = M.shape[0]
        if k == 0:
            continue
        ref = lapack_chol(M)
        got = np.tril(F) if uplo == L else np.triu(F).T
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= 1e-12, (k, err)
        assert U.scaled_residual(M, X, R) < 10 * U.EPS64, k


def test_sposv_vbatched(plasma):
    mats, rng = systems(48, 5)
    rhs = [rng.standard_normal((M.shape[0], 1)) for M in mats]
    rc, info, outA, outB = run("s", L, [M.astype(np.float32) for M in mats],
                               [R.astype(np.float32) for R in rhs], 1)
    assert rc == 0 and np.all(info == 0)
    for M, R, X in zip(mats, rhs, outB):
        res = U.scaled_residual(M.astype(np.float32).astype(np.float64),
                                X.astype(np.float64),
                                R.astype(np.float32).astype(np.float64))
        assert res < 10 * U.EPS32, (M.shape[0], res)


def test_vbatched_info(plasma):
    mats, rng = systems(0, 7, [200, 300, 64, 513 - 1])
    bad = mats[1].copy()
    bad[150, :] = 0.0
    bad[:, 150] = 0.0
    mats[1] = bad
    rhs = [rng.standard_normal((M.shape[0], 1)) for M in mats]
    rc, info, outA, outB = run("d", L, mats, rhs, 1)
    assert rc == 0
    assert list(info) == [0, 151, 0, 0]
    for j in (0, 2, 3):
        assert U.scaled_residual(mats[j], outB[j], rhs[j]) < 10 * U.EPS64


def test_dpotrf_vbatched(plasma):
    import torch
    mats, _ = systems(20, 9)
    A, _, aoff, _ = pack(mats, [np.zeros((1, 1))] * len(mats), np.float64)
    n = np.array([M.shape[0] for M in mats], dtype=np.int32)
    info = np.zeros(len(mats), dtype=np.int32)
    rc = pa.load().plasma_dpotrf_vbatched(L, len(mats), n.ctypes.data,
                                          A.data_ptr(), aoff.ctypes.data,
                                          info.ctypes.data)
    torch.cuda.synchronize()
    assert rc == 0 and np.all(info == 0)
    Ah = A.cpu().numpy()
    for o, M in zip(aoff, mats):
        F = Ah[int(o):int(o) + M.size].reshape(M.shape, order="F")
        ref = lapack_chol(M)
        assert np.linalg.norm(np.tril(F) - ref) / np.linalg.norm(ref) <= 1e-12


def test_vbatched_argument_codes(plasma):
    import torch
    lib = pa.load()
    A = torch.zeros(1000, dtype=torch.float64, device="cuda")
    n = np.array([10], dtype=np.int32)
    big = np.array([513], dtype=np.int32)
    off = np.zeros(1, dtype=np.uint64)
    info = np.zeros(1, dtype=np.int32)
    host = np.zeros(1000)
    f = lib.plasma_dposv_vbatched
    assert f(0, 1, n.ctypes.data, A.data_ptr(), off.ctypes.data, 0, None,
             None, info.ctypes.data) == -1
    assert f(L, -1, n.ctypes.data, A.data_ptr(), off.ctypes.data, 0, None,
             None, info.ctypes.data) == -2
    assert f(L, 1, big.ctypes.data, A.data_ptr(), off.ctypes.data, 0, None,
             None, info.ctypes.data) == -3
    assert f(L, 1, n.ctypes.data, host.ctypes.data, off.ctypes.data, 0, None,
             None, info.ctypes.data) == -4
    assert f(L, 1, n.ctypes.data, A.data_ptr(), off.ctypes.data, -1, None,
             None, info.ctypes.data) == -6
    assert f(L, 1, n.ctypes.data, A.data_ptr(), off.ctypes.data, 1, None,
             None, info.ctypes.data) == -7
    assert f(L, 0, n.ctypes.data, A.data_ptr(), off.ctypes.data, 0, None,
             None, info.ctypes.data) == 0
