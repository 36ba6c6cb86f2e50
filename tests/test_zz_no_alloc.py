"""plasma_dposv_vbatched_noalloc (the batched solve whose scratch is a fixed
per-device array, for callers that cannot allocate) against
plasma_dposv_vbatched through the C ABI: bit-identical factors, solutions
and info on the same systems, two calls in a row (the second with a smaller
batch, so the shared scratch holds the first call's contents), host and
device metadata arrays, a system that is not positive definite, and the
error codes (including the batch limit)."""
import numpy as np
import pytest

import plasma_amd as pa
from tests.test_gpu_batched import pack, rbf

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper
LIMIT = 65536          # PLASMA_VBATCH_NOALLOC_MAX


def _call(name, uplo, mats, rhs, nrhs, device_meta):
    import torch
    A, B, aoff, boff = pack(mats, rhs, np.float64)
    n = np.array([M.shape[0] for M in mats], dtype=np.int32)
    info = np.full(len(mats), -7, dtype=np.int32)
    keep = []
    if device_meta:
        nd = torch.from_numpy(n).cuda()
        ad = torch.from_numpy(aoff.view(np.int64)).cuda()
        bd = torch.from_numpy(boff.view(np.int64)).cuda()
        idv = torch.from_numpy(info).cuda()
        keep = [nd, ad, bd, idv]
        args = (nd.data_ptr(), A.data_ptr(), ad.data_ptr(), nrhs,
                B.data_ptr(), bd.data_ptr(), idv.data_ptr())
    else:
        args = (n.ctypes.data, A.data_ptr(), aoff.ctypes.data, nrhs,
                B.data_ptr(), boff.ctypes.data, info.ctypes.data)
    torch.cuda.synchronize()
    print("no-alloc test: %s batch=%d nrhs=%d device_meta=%s"
          % (name, len(mats), nrhs, device_meta), flush=True)
    rc = getattr(pa.load(), name)(uplo, len(mats), *args)
    torch.cuda.synchronize()
    if device_meta:
        info = keep[3].cpu().numpy()
    return rc, info, A.cpu().numpy(), B.cpu().numpy()


def _systems(count, seed, bad=None):
    rng = np.random.default_rng(seed)
    sizes = [0, 1, 2, 33, 512] + list(rng.integers(1, 513, count - 5))
    mats = [rbf(int(k), rng) for k in sizes]
    if bad is not None:
        mats[bad][3, 3] = -1.0          # leading minor 4 not positive
    rhs = [rng.standard_normal((M.shape[0], 3)) for M in mats]
    return mats, rhs


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("device_meta", [False, True])
def test_noalloc_matches_vbatched(plasma, uplo, device_meta):
    # two calls in a row, the second smaller: the scratch keeps the first
    # call's order, ticket and slabs, which the second must not depend on
    for count, seed, bad in ((300, 11, 7), (37, 12, None)):
        mats, rhs = _systems(count, seed, bad)
        ref = _call("plasma_dposv_vbatched", uplo, mats, rhs, 3, device_meta)
        got = _call("plasma_dposv_vbatched_noalloc", uplo, mats, rhs, 3,
                    device_meta)
        assert ref[0] == 0 and got[0] == 0
        assert np.array_equal(got[1], ref[1]), count
        if bad is not None:
            assert got[1][bad] == 4
            assert np.all(np.delete(got[1], bad) == 0)
        else:
            assert np.all(got[1] == 0)
        assert np.array_equal(got[2], ref[2]), count      # factors
        assert np.array_equal(got[3], ref[3]), count      # solutions


def test_noalloc_codes(plasma):
    import torch
    lib = pa.load()
    fn = lib.plasma_dposv_vbatched_noalloc
    A = torch.zeros(64, dtype=torch.float64, device="cuda")
    B = torch.zeros(64, dtype=torch.float64, device="cuda")
    n = np.array([4, 600], dtype=np.int32)
    off = np.zeros(2, dtype=np.uint64)
    info = np.zeros(2, dtype=np.int32)
    a, b = A.data_ptr(), B.data_ptr()
    assert fn(7, 2, n.ctypes.data, a, off.ctypes.data, 1, b, off.ctypes.data,
              info.ctypes.data) == -1
    assert fn(L, -1, n.ctypes.data, a, off.ctypes.data, 1, b,
              off.ctypes.data, info.ctypes.data) == -2
    assert fn(L, 0, None, a, None, 1, b, None, None) == 0     # quick return
    # an order above 512: rejected before any launch, A untouched
    A.fill_(3.0)
    torch.cuda.synchronize()
    assert fn(L, 2, n.ctypes.data, a, off.ctypes.data, 1, b,
              off.ctypes.data, info.ctypes.data) == -3
    nd = torch.from_numpy(n).cuda()
    torch.cuda.synchronize()
    assert fn(L, 2, nd.data_ptr(), a, off.ctypes.data, 1, b,
              off.ctypes.data, info.ctypes.data) == -3
    assert bool((A == 3.0).all())
    # above the scratch's capacity: -2, nothing touched
    big = LIMIT + 1
    nb = np.ones(big, dtype=np.int32)
    ob = np.zeros(big, dtype=np.uint64)
    ib = np.zeros(big, dtype=np.int32)
    assert fn(L, big, nb.ctypes.data, a, ob.ctypes.data, 1, b, ob.ctypes.data,
              ib.ctypes.data) == -2
    assert bool((A == 3.0).all())
