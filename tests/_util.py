"""Shared helpers for the tests: input generators in the reference tester's
style, the tester's error metrics, and raw device copies through HIP."""
import ctypes
import os

import numpy as np

EPS64 = np.finfo(np.float64).eps / 2   # LAPACK dlamch('E') = 2^-53
EPS32 = np.finfo(np.float32).eps / 2

_mkl = None


def mkl():
    """MKL runtime via dlopen (None when absent).  Its sequential threading
    layer is selected before any MKL routine runs: MKL keeps the layer of
    its first call, and the oracle's MKL backend needs the sequential one
    (the Intel-OpenMP layer computes wrong results next to torch's
    libgomp)."""
    global _mkl
    if _mkl is None:
        os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
        for p in ("/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so.1"):
            try:
                _mkl = ctypes.CDLL(p)
                break
            except OSError:
                continue
        else:
            _mkl = False
        if _mkl:
            _mkl.MKL_Set_Threading_Layer.restype = ctypes.c_int
            _mkl.MKL_Set_Threading_Layer(1)   # MKL_THREADING_SEQUENTIAL
    return _mkl or None


def tester_spd(n, lda=None, seed=(0, 0, 0, 1), dtype=np.float64):
    """The reference tester's SPD input (test/test_zposv.c:87-106): entries
    U(0,1) from dlarnv(idist=1, iseed), diagonal += n, upper = lower^T.
    Uses LAPACK's dlarnv from MKL when present (bit-identical to the
    tester); otherwise NumPy's uniform generator."""
    lda = lda or n
    m = mkl()
    a = np.zeros(lda * n)
    if m is not None and n > 0:
        iseed = (ctypes.c_int * 4)(*seed)
        m.LAPACKE_dlarnv(1, iseed, lda * n, a.ctypes.data_as(ctypes.c_void_p))
    else:
        a = np.random.default_rng(sum(seed)).random(lda * n)
    A = a.reshape(n, lda).T  # column-major view: A[i, j] = a[j*lda + i]
    A = np.asfortranarray(A)
    for i in range(n):
        A[i, i] += n
        A[:i, i] = A[i, :i]
    return A.astype(dtype, order="F")


def ggt_spd(n, seed=0, dtype=np.float64):
    """A = G G^T / n + I with G iid N(0,1) (the benchmark's input)."""
    g = np.random.default_rng(seed).standard_normal((n, n))
    return np.asfortranarray((g @ g.T) / n + np.eye(n)).astype(dtype)


def gp_kernel(n, d=8, lengthscale=1.0, noise=0.1, seed=0):
    """K = RBF(X, lengthscale) + noise I with X (n x d) iid N(0,1): the
    Gaussian-process log-likelihood workload of the adjoint."""
    X = np.random.default_rng(seed).standard_normal((n, d))
    sq = np.sum(X * X, axis=1)
    D2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * (X @ X.T), 0.0)
    K = np.exp(-D2 / (2.0 * lengthscale ** 2)) + noise * np.eye(n)
    return np.asfortranarray(K)


def householder_spd(n, kappa, nrefl=64, seed=0):
    """A = Q diag(s) Q^T with Q = H_1 ... H_nrefl (Householder reflectors of
    N(0,1) vectors) and s_i = kappa^(-i/(n-1)), so ||A||_2 = 1 and
    cond_2(A) = kappa: the mixed-precision benchmark's matrices, formed in
    O(n^2 nrefl) through the compact WY form Q = I - V T V^T."""
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((n, nrefl))
    s = kappa ** (-np.arange(n) / max(n - 1.0, 1.0))
    G = V.T @ V
    tau = 2.0 / np.diag(G)
    T = np.zeros((nrefl, nrefl))
    for j in range(nrefl):
        T[j, j] = tau[j]
        if j:
            T[:j, j] = -tau[j] * (T[:j, :j] @ G[:j, j])
    # A = D - U W^T - W U^T + U (V^T D V) U^T with U = V T, W = D V
    Uv = V @ T
    W = V * s[:, None]
    M = V.T @ W
    A = Uv @ (M @ Uv.T - W.T)
    A -= W @ Uv.T
    A[np.diag_indices(n)] += s
    A = (A + A.T) * 0.5
    return np.asfortranarray(A)


def rhs(n, nrhs, seed=1, dtype=np.float64):
    return np.asfortranarray(
        np.random.default_rng(seed).standard_normal((n, nrhs))).astype(dtype)


def sym_full(a, uplo_lower):
    """Symmetric matrix from one stored triangle."""
    t = np.tril(a) if uplo_lower else np.triu(a)
    return t + t.T - np.diag(np.diag(a))


def scaled_residual(A, X, B):
    """||B - A X||_inf / (n ||A||_inf ||X||_inf), the tester's criterion
    (test/test_zposv.c:144-168); A is the full symmetric matrix."""
    n = A.shape[0]
    r = B - A @ X
    an = np.abs(A).sum(axis=1).max()
    xn = np.abs(X).sum(axis=1).max()
    rn = np.abs(r).sum(axis=1).max()
    return rn / (n * an * xn)


def factor_error(F, Fref, uplo_lower):
    """||F - Fref||_F / ||Fref (stored triangle)||_F over the whole array,
    the tester's potrf criterion (test/test_zpotrf.c:124-142)."""
    tri = np.tril if uplo_lower else np.triu
    den = np.linalg.norm(tri(Fref))
    return np.linalg.norm(F - Fref) / (den if den else 1.0)


# --- raw HIP copies (device descriptors) ---------------------------------
_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"),
                                        "lib", "libamdhip64.so"))
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    return _hip


def d2h(ptr, count, dtype=np.float64):
    out = np.empty(count, dtype=dtype)
    assert hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0
    return out


def h2d(ptr, arr):
    arr = np.ascontiguousarray(arr)
    assert hip().hipMemcpy(ptr, arr.ctypes.data, arr.nbytes, 1) == 0
