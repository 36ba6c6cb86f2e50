"""The Cholesky and potrs adjoints behind torch.autograd (test harness
only; the library itself has no Python dependency).

Cholesky: forward plasma_dpotrf, backward plasma_dpotrf_adjoint.  Solve:
forward plasma_dpotrs with a given factor, backward plasma_dpotrs_adjoint
(gradients with respect to the matrix and the right-hand side).  Inputs are
symmetrised first, (K + K^T) / 2, so the symmetric gradients the adjoints
return are also the gradients with respect to each entry.  Checked: the
Gaussian-process log-likelihood gradient against the closed form
1/2 (alpha alpha^T - K^{-1}) with K^{-1} from the oracle's FP64 solve (and
the library's plasma_dpotri against it), and torch.autograd.gradcheck
(central differences, every entry) on a small kernel matrix.

Factor retention: the Cholesky function keeps its factor for the backward
pass, unless the factor would not fit the memory budget (a fraction of the
free device memory, or an explicit byte budget); then it keeps only the
input and factors it again in the backward pass.  Both ways give the same
gradient bit for bit (the factorization is deterministic)."""
import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L = pa.PlasmaLower


def _functions():
    import torch

    class Cholesky(torch.autograd.Function):
        """L = chol(A), A symmetric positive definite (lower factor).
        policy: "keep" stores the factor for the backward pass, "recompute"
        stores A and factors it again there, "auto" keeps the factor when
        it fits `budget` bytes (default: half the free device memory)."""

        @staticmethod
        def forward(ctx, A, policy="auto", budget=None):
            n = A.shape[0]
            # a row-major tensor read as a column-major array is A^T = A;
            # after potrf its column-major lower triangle is L
            F = A.detach().clone().contiguous()
            torch.cuda.synchronize()
            assert pa.potrf(L, n, F, n) == 0
            if policy == "auto":
                if budget is None:
                    budget = torch.cuda.mem_get_info()[0] // 2
                policy = "keep" if F.numel() * F.element_size() <= budget \
                    else "recompute"
            ctx.policy = policy
            out = F.t().tril()
            if policy == "keep":
                ctx.save_for_backward(F)
            else:
                ctx.save_for_backward(A.detach())
                del F
            return out

        @staticmethod
        def backward(ctx, gL):
            (S,) = ctx.saved_tensors
            n = S.shape[0]
            if ctx.policy == "keep":
                F = S
            else:
                F = S.clone().contiguous()
                torch.cuda.synchronize()
                assert pa.potrf(L, n, F, n) == 0
            G = gL.tril().t().contiguous()      # Lbar, column major
            torch.cuda.synchronize()
            assert pa.potrf_adjoint(L, n, F, n, G, n) == 0
            return G, None, None                # symmetric Abar

    class Solve(torch.autograd.Function):
        """X = A^{-1} Y from the factor F of A (column-major lower)."""

        @staticmethod
        def forward(ctx, A, F, Y):
            n = A.shape[0]
            # column major; a copy (for one column, t() of a contiguous
            # tensor is contiguous already and would alias Y)
            Xc = Y.detach().reshape(n, -1).t().clone(
                memory_format=torch.contiguous_format)
            k = Xc.shape[0]
            torch.cuda.synchronize()
            assert pa.potrs(L, n, k, F, n, Xc, n) == 0
            ctx.save_for_backward(F, Xc)
            ctx.shape = Y.shape
            return Xc.t().reshape(Y.shape)

        @staticmethod
        def backward(ctx, gX):
            F, Xc = ctx.saved_tensors
            n, k = F.shape[0], Xc.shape[0]
            Xbar = gX.reshape(n, -1).t().clone(
                memory_format=torch.contiguous_format)
            Abar = torch.zeros(n, n, dtype=F.dtype, device=F.device)
            torch.cuda.synchronize()
            assert pa.potrs_adjoint(L, n, k, F, n, Xc, n, Xbar, n, Abar,
                                    n) == 0
            # Xbar now holds Ybar = A^{-1} Xbar; Abar = -sym(Ybar X^T)
            return Abar, None, Xbar.t().reshape(ctx.shape)

    def loglik(K, y, policy="auto", budget=None):
        """GP log-likelihood -1/2 y^T K^{-1} y - 1/2 log det K (constant
        dropped) through the two functions."""
        Ks = 0.5 * (K + K.t())
        Lf = Cholesky.apply(Ks, policy, budget)
        F = Lf.detach().t().contiguous()        # the factor, column major
        alpha = Solve.apply(Ks, F, y)
        return -0.5 * (y * alpha).sum() - torch.log(torch.diagonal(Lf)).sum()

    return loglik


def test_gp_loglik_gradient_closed_form(plasma, nb_guard):
    import torch
    n = 700
    pa.set_param(pa.PlasmaNb, 128)
    Kn = U.gp_kernel(n, seed=31)
    yn = U.rhs(n, 1, seed=32)[:, 0]
    loglik = _functions()
    K = torch.from_numpy(Kn).cuda().requires_grad_(True)
    y = torch.from_numpy(yn).cuda()
    f = loglik(K, y)
    f.backward()
    G = K.grad.cpu().numpy()
    ora.set_blas(ora.MKL)
    Fo, info = ora.potrf(ora.LOWER, Kn, 128)
    assert info == 0
    Kinv, _ = ora.potrs(ora.LOWER, Fo, np.eye(n), 128)
    alpha, _ = ora.potrs(ora.LOWER, Fo, yn.reshape(n, 1), 128)
    want = 0.5 * (alpha @ alpha.T - Kinv)
    # the library's own inverse from the factor agrees with the oracle's
    Fl = np.asfortranarray(np.tril(Fo))
    assert pa.load().plasma_dpotri(L, n, Fl.ctypes.data, n) == 0
    assert np.abs(np.tril(Fl) - np.tril(Kinv)).max() <= \
        1e-10 * np.abs(Kinv).max()
    assert np.array_equal(G, G.T)
    err = np.linalg.norm(G - want) / np.linalg.norm(want)
    assert err < 1e-10, err
    # the value itself: -1/2 y^T alpha - sum log L_ii
    fv = -0.5 * float(yn @ alpha[:, 0]) - float(np.log(np.diag(Fo)).sum())
    assert abs(f.item() - fv) < 1e-9 * abs(fv)


def test_gp_loglik_gradcheck(plasma, nb_guard):
    """Central differences over every entry of K and y (n = 10)."""
    import torch
    n = 10
    K = torch.from_numpy(U.gp_kernel(n, seed=41)).cuda().requires_grad_(True)
    y = torch.from_numpy(U.rhs(n, 1, seed=42)[:, 0]).cuda() \
        .requires_grad_(True)
    loglik = _functions()
    assert torch.autograd.gradcheck(loglik, (K, y), eps=1e-6, atol=1e-6,
                                    rtol=1e-5)


def test_factor_retention_policies(plasma, nb_guard):
    """Keeping the factor or recomputing it in the backward pass gives the
    same gradient bit for bit; "auto" recomputes when the factor does not
    fit the budget."""
    import torch
    n = 900
    pa.set_param(pa.PlasmaNb, 128)
    Kn = U.gp_kernel(n, seed=51)
    yn = U.rhs(n, 1, seed=52)[:, 0]
    loglik = _functions()
    grads = {}
    for policy, budget in (("keep", None), ("recompute", None),
                           ("auto", 1024)):
        K = torch.from_numpy(Kn).cuda().requires_grad_(True)
        y = torch.from_numpy(yn).cuda()
        loglik(K, y, policy, budget).backward()
        grads[policy] = K.grad.cpu().numpy()
    assert np.array_equal(grads["keep"], grads["recompute"])
    assert np.array_equal(grads["keep"], grads["auto"])
