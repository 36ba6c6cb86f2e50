"""GPU parity tests of the norms through the C ABI: plasma_dlansy /
plasma_dlange / plasma_damax and their single precision and tile forms,
against the oracle (tests/test_oracle_norms.py pins the oracle to LAPACK).

The symmetric Inf norm is the kernel plasma_dsposv uses for ||A||_inf in its
convergence test, and the column max-norm the one it uses for ||X|| and
||R||, so these comparisons also pin those values of the mixed solver.

Max norms and row / column max-norms are exact (no rounding); sums agree
with the oracle's LAPACK-order sums to within n eps64 (different
summation order)."""
import ctypes

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

NORMS = [pa.PlasmaMaxNorm, pa.PlasmaOneNorm, pa.PlasmaInfNorm,
         pa.PlasmaFrobeniusNorm]


def _tol(norm, n):
    return 0.0 if norm == pa.PlasmaMaxNorm else 4 * n * U.EPS64


def _sym_with_junk(n, lower, seed):
    rng = np.random.default_rng(seed)
    S = rng.standard_normal((n, n)) * 10.0 ** rng.uniform(-3, 3, n)[:, None]
    S = S + S.T
    junk = np.full((n, n), np.nan)
    A = np.where(np.tril(np.ones((n, n), bool)) if lower
                 else np.triu(np.ones((n, n), bool)), S, junk)
    return S, np.asfortranarray(A)


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("uplo", [pa.PlasmaLower, pa.PlasmaUpper])
@pytest.mark.parametrize("n,nb", [(1, 64), (300, 128), (700, 256)])
def test_dlansy(plasma, nb_guard, norm, uplo, n, nb):
    pa.set_param(pa.PlasmaNb, nb)
    lower = uplo == pa.PlasmaLower
    S, A = _sym_with_junk(n, lower, seed=n + norm)
    lib = pa.load()
    v = lib.plasma_dlansy(norm, uplo, n, A.ctypes.data, n)
    ref = ora.lansy(norm, ora.LOWER if lower else ora.UPPER, A)
    assert abs(v - ref) <= _tol(norm, n) * ref, (v, ref)
    # float twin, computed in double and rounded once
    As = np.asfortranarray(A.astype(np.float32))
    vs = lib.plasma_slansy(norm, uplo, n, As.ctypes.data, n)
    refs = ora.lansy(norm, ora.LOWER if lower else ora.UPPER,
                     As.astype(np.float64))
    assert abs(vs - refs) <= (_tol(norm, n) + 2 ** -23) * refs


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("m,n,nb", [(1, 1, 64), (300, 170, 128),
                                    (170, 530, 128), (512, 512, 256)])
def test_dlange_host_and_device(plasma, nb_guard, norm, m, n, nb):
    import torch
    pa.set_param(pa.PlasmaNb, nb)
    rng = np.random.default_rng(m * n + norm)
    P = np.asfortranarray(rng.standard_normal((m + 5, n)))
    lib = pa.load()
    v = lib.plasma_dlange(norm, m, n, P.ctypes.data, m + 5)
    ref = ora.lange(norm, P[:m])
    assert abs(v - ref) <= _tol(norm, m * n) * ref
    dP = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
    torch.cuda.synchronize()
    vd = lib.plasma_dlange(norm, m, n, dP.data_ptr(), m + 5)
    assert vd == v      # same kernels, same order


@pytest.mark.parametrize("colrow", [pa.PlasmaColumnwise, pa.PlasmaRowwise])
@pytest.mark.parametrize("m,n,nb", [(1, 1, 64), (513, 64, 128),
                                    (100, 700, 256)])
def test_damax(plasma, nb_guard, colrow, m, n, nb):
    pa.set_param(pa.PlasmaNb, nb)
    rng = np.random.default_rng(m + n)
    P = np.asfortranarray(rng.standard_normal((m + 2, n)) * 1e3)
    out = np.full(n if colrow == pa.PlasmaColumnwise else m, -1.0)
    assert pa.load().plasma_damax(colrow, m, n, P.ctypes.data, m + 2,
                                  out.ctypes.data) == 0
    ref = ora.amax(ora.COLUMNWISE if colrow == pa.PlasmaColumnwise
                   else ora.ROWWISE, P[:m])
    assert np.array_equal(out, ref)
    Ps = np.asfortranarray(P.astype(np.float32))
    outs = np.zeros(out.shape, np.float32)
    assert pa.load().plasma_samax(colrow, m, n, Ps.ctypes.data, m + 2,
                                  outs.ctypes.data) == 0
    assert np.array_equal(outs, np.abs(Ps[:m]).max(
        axis=0 if colrow == pa.PlasmaColumnwise else 1))


def test_norm_nan_and_argument_codes(plasma):
    lib = pa.load()
    A = np.asfortranarray(np.eye(40))
    A[7, 3] = np.nan
    for norm in NORMS:
        assert np.isnan(lib.plasma_dlange(norm, 40, 40, A.ctypes.data, 40))
    a = np.eye(4, order="F")
    assert lib.plasma_dlansy(999, pa.PlasmaLower, 4, a.ctypes.data, 4) == -1
    assert lib.plasma_dlansy(pa.PlasmaMaxNorm, 999, 4, a.ctypes.data, 4) == -2
    assert lib.plasma_dlansy(pa.PlasmaMaxNorm, pa.PlasmaLower, -1,
                             a.ctypes.data, 4) == -3
    assert lib.plasma_dlansy(pa.PlasmaMaxNorm, pa.PlasmaLower, 4,
                             a.ctypes.data, 3) == -5
    assert lib.plasma_dlange(pa.PlasmaMaxNorm, -1, 4, a.ctypes.data, 4) == -2
    assert lib.plasma_dlange(pa.PlasmaMaxNorm, 4, -1, a.ctypes.data, 4) == -3
    assert lib.plasma_dlange(pa.PlasmaMaxNorm, 4, 4, a.ctypes.data, 3) == -5
    assert lib.plasma_dlange(pa.PlasmaMaxNorm, 0, 4, a.ctypes.data, 4) == 0.0
    v = np.zeros(4)
    assert lib.plasma_damax(7, 4, 4, a.ctypes.data, 4, v.ctypes.data) == -1


def test_tile_forms_on_caller_descriptors(plasma, nb_guard):
    """plasma_omp_dlansy / dlange / damax on descriptors the caller creates
    and fills (device storage), results in host and device memory, read
    after plasma_sequence_sync."""
    import torch
    lib = pa.load()
    n, nb = 600, 128
    pa.set_param(pa.PlasmaNb, nb)
    rng = np.random.default_rng(1)
    G = np.asfortranarray(rng.standard_normal((n, n)))
    A = pa.plasma_desc_t()
    assert lib.plasma_desc_general_create(pa.PlasmaRealDouble, nb, nb, n, n,
                                          0, 0, n, n, ctypes.byref(A)) == 0
    seq, req = pa.plasma_sequence_t(), pa.plasma_request_t()
    lib.plasma_sequence_init(ctypes.byref(seq))
    lib.plasma_request_init(ctypes.byref(req))
    lib.plasma_omp_dge2desc(G.ctypes.data, n, A, ctypes.byref(seq),
                            ctypes.byref(req))
    host = (ctypes.c_double * 3)()
    dev = torch.zeros(n, dtype=torch.float64, device="cuda")
    lib.plasma_omp_dlange(pa.PlasmaFrobeniusNorm, A, None,
                          ctypes.addressof(host), ctypes.byref(seq),
                          ctypes.byref(req))
    lib.plasma_omp_dlansy(pa.PlasmaInfNorm, pa.PlasmaLower, A, None,
                          ctypes.addressof(host) + 8, ctypes.byref(seq),
                          ctypes.byref(req))
    lib.plasma_omp_damax(pa.PlasmaColumnwise, A, None, dev.data_ptr(),
                         ctypes.byref(seq), ctypes.byref(req))
    assert lib.plasma_sequence_sync(ctypes.byref(seq)) == 0
    assert abs(host[0] - ora.lange(ora.FRONORM, G)) <= 1e-12 * host[0]
    assert abs(host[1] - ora.lansy(ora.INFNORM, ora.LOWER, G)) <= \
        1e-12 * host[1]
    assert np.array_equal(dev.cpu().numpy(), ora.amax(ora.COLUMNWISE, G))
    # a failed sequence stops later work: bad norm value
    lib.plasma_omp_dlange(12345, A, None, ctypes.addressof(host),
                          ctypes.byref(seq), ctypes.byref(req))
    assert seq.status == pa.PlasmaErrorIllegalValue
    lib.plasma_desc_destroy(ctypes.byref(A))
