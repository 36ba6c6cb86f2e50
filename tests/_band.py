"""Band matrices for the band Cholesky tests: SPD band matrices and their
LAPACK band arrays (column j of AB holds A(i, j) at row kd + i - j for upper
storage, i - j for lower; 0-based)."""
import numpy as np


def spd_band(n, kd, seed=0):
    """Symmetric, diagonally dominant band matrix of bandwidth kd."""
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for d in range(1, kd + 1):
        v = rng.standard_normal(n - d)
        A[np.arange(d, n), np.arange(n - d)] = v
    A = A + A.T
    A[np.diag_indices(n)] = 2.0 * kd + 1.0 + rng.random(n)
    return A


def to_band(A, kd, lower, ldab=None):
    n = A.shape[0]
    ab = np.zeros((ldab or kd + 1, n), order="F")
    for j in range(n):
        if lower:
            for i in range(j, min(n, j + kd + 1)):
                ab[i - j, j] = A[i, j]
        else:
            for i in range(max(0, j - kd), j + 1):
                ab[kd + i - j, j] = A[i, j]
    return ab


def from_band(ab, kd, lower):
    """The triangular matrix (factor) held in a band array."""
    n = ab.shape[1]
    F = np.zeros((n, n))
    for j in range(n):
        if lower:
            for i in range(j, min(n, j + kd + 1)):
                F[i, j] = ab[i - j, j]
        else:
            for i in range(max(0, j - kd), j + 1):
                F[i, j] = ab[kd + i - j, j]
    return F
