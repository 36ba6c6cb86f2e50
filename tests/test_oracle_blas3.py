"""Pins the oracle's tile BLAS-3 restatements (reference compute/pzgemm.c,
pzsyrk.c, pztrsm.c) on the CPU:

  * against SciPy's BLAS (OpenBLAS) dgemm / dsyrk / dtrsm on the same
    inputs, every transposition, side, triangle and diagonal kind, with
    edge tiles (sizes not multiples of nb), alpha != 1 and beta in {0, 1/2};
  * against exact integer arithmetic (integer matrices whose products and
    unit-triangular solves are exact in FP64, so the tile sums must agree
    to the last bit with NumPy's integer result);
  * against what the mathematics fixes: trsm with A = I returns alpha B,
    a unit-diagonal solve ignores the stored diagonal, op(A) X = B holds for
    the solution, and syrk leaves the other triangle untouched;
  * with both tile-kernel backends (built-in loops and MKL) where MKL loads.

A transposed operand, a wrong side or a swapped triangle fails the SciPy
comparison (the inputs are non-symmetric and the solves non-commuting)."""
import itertools

import numpy as np
import pytest
import scipy.linalg.blas as sb

import oracle as ora

BACKENDS = [ora.BUILTIN, ora.MKL]


def _backend(b):
    got = ora.set_blas(b)
    if got != b:
        pytest.skip("MKL not available")


def _tri(n, lower, rng):
    """Triangular, well conditioned with its diagonal or with a unit one
    (off-diagonal entries N(0, 1/n))."""
    a = rng.standard_normal((n, n)) / np.sqrt(n)
    a = np.tril(a) if lower else np.triu(a)
    a[np.diag_indices(n)] = 2.0 + rng.random(n)      # well conditioned
    return a


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("ta,tb", list(itertools.product([ora.NOTRANS, ora.TRANS],
                                                          repeat=2)))
def test_gemm_vs_scipy(backend, ta, tb):
    _backend(backend)
    rng = np.random.default_rng(ta * 7 + tb)
    m, n, k, nb = 70, 45, 53, 16
    A = rng.standard_normal((m, k) if ta == ora.NOTRANS else (k, m))
    B = rng.standard_normal((k, n) if tb == ora.NOTRANS else (n, k))
    for beta in (0.0, 0.5):
        C = rng.standard_normal((m, n))
        got = ora.gemm(ta, tb, 1.25, A, B, beta, C, nb)
        ref = sb.dgemm(1.25, A, B, beta, C.copy(order="F"),
                       trans_a=int(ta == ora.TRANS), trans_b=int(tb == ora.TRANS))
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_gemm_exact_integers_and_alpha_zero():
    ora.set_blas(ora.BUILTIN)
    rng = np.random.default_rng(3)
    A = rng.integers(-8, 9, (37, 29)).astype(float)
    B = rng.integers(-8, 9, (29, 41)).astype(float)
    C = rng.integers(-8, 9, (37, 41)).astype(float)
    got = ora.gemm(ora.NOTRANS, ora.NOTRANS, 2.0, A, B, -1.0, C, 8)
    assert np.array_equal(got, 2 * (A.astype(np.int64) @ B.astype(np.int64)) - C)
    # alpha = 0: C = beta C, operands not read
    got = ora.gemm(ora.NOTRANS, ora.NOTRANS, 0.0, np.full((37, 29), np.nan),
                   B, 3.0, C, 8)
    assert np.array_equal(got, 3 * C)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("uplo", [ora.LOWER, ora.UPPER])
@pytest.mark.parametrize("trans", [ora.NOTRANS, ora.TRANS])
def test_syrk_vs_scipy(backend, uplo, trans):
    _backend(backend)
    rng = np.random.default_rng(uplo + trans)
    n, k, nb = 61, 38, 16
    A = rng.standard_normal((n, k) if trans == ora.NOTRANS else (k, n))
    C = rng.standard_normal((n, n))
    got = ora.syrk(uplo, trans, 0.75, A, 0.5, C, nb)
    ref = sb.dsyrk(0.75, A, 0.5, C.copy(order="F"),
                   trans=int(trans == ora.TRANS), lower=int(uplo == ora.LOWER))
    tri = np.tril if uplo == ora.LOWER else np.triu
    other = np.triu if uplo == ora.LOWER else np.tril
    assert np.abs(tri(got) - tri(ref)).max() <= 1e-13 * np.abs(ref).max()
    # the other triangle (strictly) is untouched
    assert np.array_equal(other(got, 1 if uplo == ora.LOWER else -1),
                          other(C, 1 if uplo == ora.LOWER else -1))


CASES = list(itertools.product([ora.LEFT, ora.RIGHT], [ora.LOWER, ora.UPPER],
                               [ora.NOTRANS, ora.TRANS],
                               [ora.NONUNIT, ora.UNIT]))


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("side,uplo,trans,diag", CASES)
def test_trsm_vs_scipy(backend, side, uplo, trans, diag):
    _backend(backend)
    rng = np.random.default_rng(side + uplo + trans + diag)
    m, n, nb = 67, 43, 16
    na = m if side == ora.LEFT else n
    A = _tri(na, uplo == ora.LOWER, rng)
    # the other triangle holds garbage the solve must not read
    A += (np.triu if uplo == ora.LOWER else np.tril)(
        rng.standard_normal((na, na)), 1 if uplo == ora.LOWER else -1) * 100
    B = rng.standard_normal((m, n))
    got = ora.trsm(side, uplo, trans, diag, 1.5, A, B, nb)
    ref = sb.dtrsm(1.5, A, B.copy(order="F"), side=int(side == ora.RIGHT),
                   lower=int(uplo == ora.LOWER),
                   trans_a=int(trans == ora.TRANS), diag=int(diag == ora.UNIT))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # op(A) X = alpha B (Left) or X op(A) = alpha B (Right)
    T = (np.tril if uplo == ora.LOWER else np.triu)(A)
    if diag == ora.UNIT:
        T[np.diag_indices(na)] = 1.0
    opA = T.T if trans == ora.TRANS else T
    lhs = opA @ got if side == ora.LEFT else got @ opA
    scale = (np.abs(opA) @ np.abs(got) if side == ora.LEFT
             else np.abs(got) @ np.abs(opA))
    assert np.all(np.abs(lhs - 1.5 * B) <= 4 * na * 2.0 ** -53 * scale)


def test_trsm_identity_and_exact_unit_solve():
    ora.set_blas(ora.BUILTIN)
    rng = np.random.default_rng(9)
    B = rng.standard_normal((50, 30))
    got = ora.trsm(ora.LEFT, ora.LOWER, ora.NOTRANS, ora.NONUNIT, -2.0,
                   np.eye(50), B, 16)
    assert np.array_equal(got, -2.0 * B)
    # unit lower L with small integers: L X = B with X integer is exact;
    # Right / Upper / Trans with A = L^T: op(A) = L, so Y Lu = X^T Lu
    L = np.tril(rng.integers(-2, 3, (40, 40)), -1).astype(float)
    L[np.diag_indices(40)] = 7.0           # must be ignored (unit)
    Lu = L.copy()
    Lu[np.diag_indices(40)] = 1.0
    X = rng.integers(-3, 4, (40, 5)).astype(float)
    B = Lu @ X
    got = ora.trsm(ora.LEFT, ora.LOWER, ora.NOTRANS, ora.UNIT, 1.0, L, B, 8)
    assert np.array_equal(got, X)
    got = ora.trsm(ora.RIGHT, ora.UPPER, ora.TRANS, ora.UNIT, 1.0, L.T,
                   X.T @ Lu, 8)
    assert np.array_equal(got, X.T)


def test_trsm_single_precision():
    ora.set_blas(ora.BUILTIN)
    rng = np.random.default_rng(5)
    A = _tri(48, True, rng)
    B = rng.standard_normal((48, 20))
    got = ora.trsm(ora.LEFT, ora.LOWER, ora.TRANS, ora.NONUNIT, 1.0, A, B, 16,
                   dtype=np.float32)
    ref = np.linalg.solve(np.tril(A).T, B)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
