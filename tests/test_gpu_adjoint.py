"""GPU parity tests of the Cholesky and potrs adjoints through the C ABI
(no reference counterpart).  The oracle's adjoint is the reverse-mode sweep
of the scalar Cholesky, a different algorithm from the device's closed
form, pinned in test_oracle_adjoint.py by finite differences and closed
forms.  Checks: device vs oracle on Gaussian-process kernels (both
triangles, edge tiles, the unread triangle filled with NaN), FP32 against
the FP64 oracle, the log-determinant closed form Abar = A^{-1} at a larger
order, the potrs adjoint against its composition from oracle solves, the
GP log-likelihood gradient 1/2 (a a^T - K^{-1}) against the closed form
and central finite differences (h = 1e-6) of the device's own forward
pass, device-pointer arguments and the argument codes."""
import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L, Up = pa.PlasmaLower, pa.PlasmaUpper


def _ora_uplo(uplo):
    return ora.LOWER if uplo == L else ora.UPPER


def _factor(uplo, K, nb=64):
    F, info = ora.potrf(_ora_uplo(uplo), K, nb)
    assert info == 0
    return F


def _tri(uplo, M):
    return np.tril(M) if uplo == L else np.triu(M)


def _with_nan(uplo, T):
    """T in the uplo triangle, NaN in the other (never read)."""
    out = np.full(T.shape, np.nan, order="F")
    mask = np.tril(np.ones(T.shape, bool)) if uplo == L \
        else np.triu(np.ones(T.shape, bool))
    out[mask] = T[mask]
    return out


def _relerr(X, Y):
    return np.linalg.norm(X - Y) / np.linalg.norm(Y)


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nb", [(1, 64), (100, 64), (256, 64), (515, 128),
                                  (700, 256)])
def test_potrf_adjoint_vs_oracle(plasma, nb_guard, uplo, n, nb):
    pa.set_param(pa.PlasmaNb, nb)
    K = U.gp_kernel(n, seed=n)
    F = _factor(uplo, K)
    Lbar = _tri(uplo, np.random.default_rng(n + 1).standard_normal((n, n)))
    want, info = ora.potrf_adjoint(_ora_uplo(uplo), F, Lbar)
    assert info == 0
    Fa = np.asfortranarray(F)
    G = _with_nan(uplo, Lbar)
    assert pa.potrf_adjoint(uplo, n, Fa, n, G, n) == 0
    assert np.array_equal(G, G.T)
    assert _relerr(G, want) < 1e-12, _relerr(G, want)


@pytest.mark.parametrize("uplo", [L, Up])
def test_spotrf_adjoint_vs_oracle(plasma, nb_guard, uplo):
    n = 600
    pa.set_param(pa.PlasmaNb, 128)
    K = U.gp_kernel(n, seed=11)
    F = _factor(uplo, K)
    Lbar = _tri(uplo, np.random.default_rng(12).standard_normal((n, n)))
    want, _ = ora.potrf_adjoint(_ora_uplo(uplo), F, Lbar)
    Fs = np.asfortranarray(F.astype(np.float32))
    G = np.asfortranarray(_with_nan(uplo, Lbar).astype(np.float32))
    assert pa.potrf_adjoint(uplo, n, Fs, n, G, n, prec="s") == 0
    assert _relerr(G.astype(np.float64), want) < 2e-5


def test_logdet_closed_form_large(plasma, nb_guard):
    """Lbar = diag(2 / L_ii) (the adjoint of log det A) gives A^{-1}."""
    import torch
    n = 4096
    pa.set_param(pa.PlasmaNb, 256)
    K = U.gp_kernel(n, seed=5)
    dK = torch.from_numpy(K).cuda()
    torch.cuda.synchronize()
    assert pa.potrf(L, n, dK, n) == 0          # device factor in place
    F = dK.cpu().numpy()
    G = torch.from_numpy(np.asfortranarray(np.diag(2.0 / np.diag(F)))).cuda()
    torch.cuda.synchronize()
    assert pa.potrf_adjoint(L, n, dK, n, G, n) == 0
    Gh = G.cpu().numpy()
    R = Gh @ K - np.eye(n)
    assert np.max(np.abs(R)) < 1e-11, np.max(np.abs(R))


@pytest.mark.parametrize("uplo", [L, Up])
@pytest.mark.parametrize("n,nrhs,nb", [(300, 3, 64), (513, 17, 128)])
def test_potrs_adjoint_vs_composition(plasma, nb_guard, uplo, n, nrhs, nb):
    pa.set_param(pa.PlasmaNb, nb)
    rng = np.random.default_rng(n)
    K = U.gp_kernel(n, seed=n + 2)
    F = _factor(uplo, K)
    X = np.asfortranarray(rng.standard_normal((n, nrhs)))
    Xbar = np.asfortranarray(rng.standard_normal((n, nrhs)))
    S = rng.standard_normal((n, n))
    A0 = np.asfortranarray(S + S.T)
    Bbar, _ = ora.potrs(_ora_uplo(uplo), F, Xbar, 64)
    want = A0 - 0.5 * (Bbar @ X.T + X @ Bbar.T)
    xb, ab = Xbar.copy(order="F"), A0.copy(order="F")
    assert pa.potrs_adjoint(uplo, n, nrhs, np.asfortranarray(F), n, X, n,
                            xb, n, ab, n) == 0
    assert _relerr(xb, Bbar) < 1e-12
    assert _relerr(ab, want) < 1e-12


def _loglik(K, y):
    """-1/2 y^T K^{-1} y - 1/2 log det K through the device forward pass."""
    n = K.shape[0]
    F = np.asfortranarray(K.copy())
    assert pa.potrf(L, n, F, n) == 0
    a = np.asfortranarray(y.copy())
    assert pa.potrs(L, n, 1, F, n, a, n) == 0
    return -0.5 * float(y[:, 0] @ a[:, 0]) - np.sum(np.log(np.diag(F))), F, a


def _gp_gradient(K, y, F, alpha):
    n = K.shape[0]
    G = np.asfortranarray(np.diag(-1.0 / np.diag(F)))
    assert pa.potrf_adjoint(L, n, F, n, G, n) == 0
    xb = np.asfortranarray(-0.5 * y)
    assert pa.potrs_adjoint(L, n, 1, F, n, alpha, n, xb, n, G, n) == 0
    return G


def test_gp_gradient_closed_form(plasma, nb_guard):
    n = 2048
    pa.set_param(pa.PlasmaNb, 256)
    K = U.gp_kernel(n, seed=21)
    y = U.rhs(n, 1, seed=22)
    _, F, alpha = _loglik(K, y)
    G = _gp_gradient(K, y, F, alpha)
    ora.set_blas(ora.MKL)
    Kinv, _ = ora.potrs(ora.LOWER, F, np.eye(n), 256)
    want = 0.5 * (alpha @ alpha.T - Kinv)
    assert _relerr(G, want) < 1e-11, _relerr(G, want)


def test_gp_gradient_finite_differences(plasma, nb_guard):
    n, h = 200, 1e-6
    pa.set_param(pa.PlasmaNb, 64)
    K = U.gp_kernel(n, seed=31)
    y = U.rhs(n, 1, seed=32)
    _, F, alpha = _loglik(K, y)
    G = _gp_gradient(K, y, F, alpha)
    scale = np.max(np.abs(G))
    rng = np.random.default_rng(33)
    pairs = [(i, i) for i in (0, 77, 199)] + \
        [tuple(sorted(rng.choice(n, 2, replace=False), reverse=True))
         for _ in range(8)]
    for i, j in pairs:
        E = np.zeros((n, n))
        E[i, j] = E[j, i] = h
        fd = (_loglik(K + E, y)[0] - _loglik(K - E, y)[0]) / (2 * h)
        want = G[i, j] if i == j else 2.0 * G[i, j]
        assert abs(fd - want) <= 1e-5 * scale, (i, j, fd, want)


def test_adjoint_argument_codes(plasma):
    a = np.eye(4, order="F")
    g = np.zeros((4, 4), order="F")
    x = np.zeros((4, 2), order="F")
    assert pa.potrf_adjoint(999, 4, a, 4, g, 4) == -1
    assert pa.potrf_adjoint(L, -1, a, 4, g, 4) == -2
    assert pa.potrf_adjoint(L, 4, a, 3, g, 4) == -4
    assert pa.potrf_adjoint(L, 4, a, 4, g, 3) == -6
    assert pa.potrf_adjoint(L, 0, a, 4, g, 4) == 0
    assert pa.potrs_adjoint(999, 4, 2, a, 4, x, 4, x, 4, g, 4) == -1
    assert pa.potrs_adjoint(L, -1, 2, a, 4, x, 4, x, 4, g, 4) == -2
    assert pa.potrs_adjoint(L, 4, -1, a, 4, x, 4, x, 4, g, 4) == -3
    assert pa.potrs_adjoint(L, 4, 2, a, 3, x, 4, x, 4, g, 4) == -5
    assert pa.potrs_adjoint(L, 4, 2, a, 4, x, 3, x, 4, g, 4) == -7
    assert pa.potrs_adjoint(L, 4, 2, a, 4, x, 4, x, 3, g, 4) == -9
    assert pa.potrs_adjoint(L, 4, 2, a, 4, x, 4, x, 4, g, 3) == -11
    assert pa.potrs_adjoint(L, 4, 0, a, 4, x, 4, x, 4, g, 4) == 0
