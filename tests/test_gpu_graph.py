"""hipGraph mode of the factorization (PlasmaGraphCapture): the launch
sequence is captured on the first call and replayed on later calls with the
same descriptor.  Every replay is checked against the oracle, with eager
(uncaptured) work and synchronisations between replays; a failed pivot must
still come back as LAPACK's info, and the next replay must start clean."""
import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper


@pytest.fixture
def graph_mode(plasma, nb_guard):
    pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaEnabled)
    yield
    pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaDisabled)


def _device(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("lookahead", [0, 1, 3])
def test_graph_replays(graph_mode, uplo, lookahead):
    import torch
    n, nb = 1000, 128
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    pa.set_param(pa.PlasmaLookahead, lookahead)
    ora.set_blas(ora.MKL)
    for rep in range(4):
        A = U.ggt_spd(n, seed=40 + rep)
        da = _device(A)
        torch.cuda.synchronize()
        assert pa.potrf(uplo, n, da, n) == 0
        F = da.cpu().numpy().T
        ref, rinfo = ora.potrf(uplo, A, nb)
        assert rinfo == 0
        assert U.factor_error(F, ref, uplo == L) <= 1e-12, rep
        # eager work between replays
        pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaDisabled)
        B = U.tester_spd(300, 300)
        b = B.copy(order="F")
        assert pa.potrf(L, 300, b, 300) == 0
        pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaEnabled)


def test_graph_info_then_clean(graph_mode):
    import torch
    n, nb = 768, 128
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    A = U.ggt_spd(n, seed=50)
    bad = A.copy()
    bad[300, :] = 0.0
    bad[:, 300] = 0.0
    # capture on a good matrix, then a failing replay, then a good one
    for M, want in ((A, 0), (bad, 301), (A, 0)):
        da = _device(M)
        torch.cuda.synchronize()
        assert pa.potrf(L, n, da, n) == want


def test_graph_dposv(graph_mode):
    import torch
    n, nb, nrhs = 2048, 256, 16
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, nb)
    for rep in range(3):
        A = U.ggt_spd(n, seed=60 + rep)
        B = U.rhs(n, nrhs, seed=70 + rep)
        da, db = _device(A), _device(B)
        torch.cuda.synchronize()
        assert pa.posv(L, n, nrhs, da, n, db, n) == 0
        X = db.cpu().numpy().T
        assert U.scaled_residual(A, X, B) < 10 * U.EPS64
