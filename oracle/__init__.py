"""Python access to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

# MKL's tile kernels run sequentially inside the oracle's OpenMP tasks.
# MKL fixes its threading layer at its first routine call in the process,
# so the choice is also made through the environment, before anything
# (a test helper, a generator) can call MKL first.
os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LOWER, UPPER = 122, 121
BUILTIN, MKL = 0, 1

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-B", "-s"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        i, p, d = ctypes.c_int, ctypes.c_void_p, ctypes.c_double
        f = ctypes.c_float
        pi = ctypes.POINTER(ctypes.c_int)
        sig = {
            "ora_set_blas": (i, [i]), "ora_get_blas": (i, []),
            "ora_mkl_refused": (i, []),
            "ora_dpbtrf": (i, [i, i, i, p, i, i]),
            "ora_spbtrf": (i, [i, i, i, p, i, i]),
            "ora_dpbsv": (i, [i, i, i, i, p, i, p, i, i, i]),
            "ora_spbsv": (i, [i, i, i, i, p, i, p, i, i, i]),
            "ora_dpotri": (i, [i, i, p, i, i]),
            "ora_dpoinv": (i, [i, i, p, i, i]),
            "ora_spotri": (i, [i, i, p, i, i]),
            "ora_dlansy": (d, [i, i, i, p, i]),
            "ora_dlange": (d, [i, i, i, p, i]),
            "ora_damax": (None, [i, i, i, p, i, p]),
            "ora_set_threads": (None, [i]), "ora_get_threads": (i, []),
            "ora_plant_mistake": (None, [i]),
            "ora_dpotrf": (i, [i, i, p, i, i]),
            "ora_spotrf": (i, [i, i, p, i, i]),
            "ora_dpotrs": (i, [i, i, i, p, i, p, i, i]),
            "ora_dposv": (i, [i, i, i, p, i, p, i, i]),
            "ora_sposv": (i, [i, i, i, p, i, p, i, i]),
            "ora_dsposv": (i, [i, i, i, p, i, p, i, p, i, pi, i]),
            "ora_dlag2s": (i, [i, i, p, i, p, i]),
            "ora_slag2d": (None, [i, i, p, i, p, i]),
            "ora_dpotrf_adjoint": (i, [i, i, p, i, p, i]),
            "ora_tile_storage_elems": (ctypes.c_long, [i, i, i, i]),
            "ora_lapack2tile": (None, [i, i, i, p, i, p, i]),
            "ora_tile2lapack": (None, [i, i, i, p, p, i, i]),
            "ora_dgemm": (i, [i, i, i, i, i, d, p, i, p, i, d, p, i, i]),
            "ora_dsyrk": (i, [i, i, i, i, d, p, i, d, p, i, i]),
            "ora_dtrsm": (i, [i, i, i, i, i, i, d, p, i, p, i, i]),
            "ora_sgemm": (i, [i, i, i, i, i, f, p, i, p, i, f, p, i, i]),
            "ora_strsm": (i, [i, i, i, i, i, i, f, p, i, p, i, i]),
            "ora_zpotrf": (i, [i, i, p, i, i]),
            "ora_cpotrf": (i, [i, i, p, i, i]),
            "ora_zposv": (i, [i, i, i, p, i, p, i, i]),
            "ora_cposv": (i, [i, i, i, p, i, p, i, i]),
            "ora_zpotrs": (i, [i, i, i, p, i, p, i, i]),
            "ora_cpotrs": (i, [i, i, i, p, i, p, i, i]),
            "ora_zcposv": (i, [i, i, i, p, i, p, i, p, i, pi, i]),
            "ora_zlag2c": (i, [i, i, p, i, p, i]),
            "ora_clag2z": (None, [i, i, p, i, p, i]),
            "ora_zlanhe_inf": (d, [i, i, p, i]),
            "ora_dzamax": (None, [i, i, p, i, p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


def _f(a, dtype):
    a = np.asarray(a, dtype=dtype)
    return a if a.flags.f_contiguous else np.asfortranarray(a)


def _ld(a):
    """Leading dimension of a column-major 2-D array (or a 1-D column)."""
    if a.ndim == 1:
        return max(1, a.shape[0])
    return max(1, a.strides[1] // a.itemsize)


def set_blas(backend):
    """Select the tile-kernel backend; returns the one in force (BUILTIN
    when MKL is absent).  Raises when MKL is present but running on a
    threaded layer, whose results are wrong in a process that also holds
    torch's libgomp."""
    got = load().ora_set_blas(backend)
    if backend == MKL and got != MKL and load().ora_mkl_refused():
        raise RuntimeError("oracle: MKL is not on its sequential threading "
                           "layer (an MKL routine ran first)")
    return got


def set_threads(n):
    load().ora_set_threads(n)


def potrf(uplo, a, nb, dtype=np.float64):
    """Returns (factor array, info); the input is not modified."""
    a = _f(a, dtype).copy(order="F")
    fn = load().ora_dpotrf if dtype == np.float64 else load().ora_spotrf
    info = fn(uplo, a.shape[0], a.ctypes.data, _ld(a), nb)
    return a, info


def posv(uplo, a, b, nb, dtype=np.float64):
    a = _f(a, dtype).copy(order="F")
    b = _f(b, dtype).copy(order="F")
    fn = load().ora_dposv if dtype == np.float64 else load().ora_sposv
    info = fn(uplo, a.shape[0], b.shape[1], a.ctypes.data, _ld(a),
              b.ctypes.data, _ld(b), nb)
    return a, b, info


def pbtrf(uplo, ab, kd, nb, dtype=np.float64):
    """Band Cholesky of the LAPACK band array ab ((ldab >= kd + 1) x n, the
    uplo triangle's kd + 1 diagonals); returns (factor band array, info),
    the input is not modified (reference compute/zpbtrf.c)."""
    ab = _f(ab, dtype).copy(order="F")
    fn = load().ora_dpbtrf if dtype == np.float64 else load().ora_spbtrf
    info = fn(uplo, ab.shape[1], kd, ab.ctypes.data, _ld(ab), nb)
    return ab, info


def pbsv(uplo, ab, kd, b, nb, factor=True, dtype=np.float64):
    """pbsv (factor=True: returns (factor, X, info)) or pbtrs (factor=False:
    ab holds the factor; returns (ab, X, info)), reference zpbsv.c /
    zpbtrs.c."""
    ab = _f(ab, dtype).copy(order="F")
    b = _f(b, dtype).copy(order="F")
    fn = load().ora_dpbsv if dtype == np.float64 else load().ora_spbsv
    info = fn(uplo, ab.shape[1], kd, b.shape[1], ab.ctypes.data, _ld(ab),
              b.ctypes.data, _ld(b), nb, 1 if factor else 0)
    return ab, b, info


def potri(uplo, f, nb, dtype=np.float64):
    """A^{-1} in the uplo triangle from the Cholesky factor f (uplo
    triangle; the other triangle passes through); (array, info)."""
    a = _f(f, dtype).copy(order="F")
    fn = load().ora_dpotri if dtype == np.float64 else load().ora_spotri
    info = fn(uplo, a.shape[0], a.ctypes.data, _ld(a), nb)
    return a, info


def poinv(uplo, a, nb):
    """A^{-1} in the uplo triangle from SPD A (factor, then potri)."""
    a = _f(a, np.float64).copy(order="F")
    info = load().ora_dpoinv(uplo, a.shape[0], a.ctypes.data, _ld(a), nb)
    return a, info


def potrs(uplo, l, b, nb):
    l = _f(l, np.float64)
    b = _f(b, np.float64).copy(order="F")
    info = load().ora_dpotrs(uplo, l.shape[0], b.shape[1], l.ctypes.data,
                             _ld(l), b.ctypes.data, _ld(b), nb)
    return b, info


def dsposv(uplo, a, b, nb):
    a = _f(a, np.float64)
    b = _f(b, np.float64)
    x = np.zeros_like(b, order="F")
    it = ctypes.c_int(0)
    info = load().ora_dsposv(uplo, a.shape[0], b.shape[1], a.ctypes.data,
                             _ld(a), b.ctypes.data, _ld(b), x.ctypes.data,
                             _ld(x), ctypes.byref(it), nb)
    return x, it.value, info


def potrf_adjoint(uplo, l, lbar):
    """Symmetric gradient Abar (full n x n) from the factor l and the
    factor's adjoint lbar, both in the uplo triangle."""
    l = _f(l, np.float64)
    g = _f(lbar, np.float64).copy(order="F")
    info = load().ora_dpotrf_adjoint(uplo, l.shape[0], l.ctypes.data, _ld(l),
                                     g.ctypes.data, _ld(g))
    return g, info


def dlag2s(a):
    a = _f(a, np.float64)
    s = np.zeros(a.shape, dtype=np.float32, order="F")
    info = load().ora_dlag2s(a.shape[0], a.shape[1], a.ctypes.data, _ld(a),
                             s.ctypes.data, _ld(s))
    return s, info


def slag2d(s):
    s = _f(s, np.float32)
    a = np.zeros(s.shape, dtype=np.float64, order="F")
    load().ora_slag2d(s.shape[0], s.shape[1], s.ctypes.data, _ld(s),
                      a.ctypes.data, _ld(a))
    return a


def lapack2tile(uplo, a, nb):
    a = _f(a, np.float64)
    m, n = a.shape
    t = np.zeros(load().ora_tile_storage_elems(uplo, m, n, nb))
    load().ora_lapack2tile(uplo, m, n, a.ctypes.data, _ld(a), t.ctypes.data, nb)
    return t


def tile2lapack(uplo, tiles, m, n, nb, init=None):
    a = np.zeros((m, n), order="F") if init is None else _f(init, np.float64).copy(order="F")
    t = np.ascontiguousarray(tiles, dtype=np.float64)
    load().ora_tile2lapack(uplo, m, n, t.ctypes.data, a.ctypes.data, _ld(a), nb)
    return a


MAXNORM, ONENORM, INFNORM, FRONORM = 177, 171, 175, 174
COLUMNWISE, ROWWISE = 401, 402


def lansy(norm, uplo, a):
    a = _f(a, np.float64)
    return load().ora_dlansy(norm, uplo, a.shape[0], a.ctypes.data, _ld(a))


def lange(norm, a):
    a = _f(a, np.float64)
    return load().ora_dlange(norm, a.shape[0], a.shape[1], a.ctypes.data,
                             _ld(a))


def amax(colrow, a):
    a = _f(a, np.float64)
    m, n = a.shape
    v = np.zeros(n if colrow == COLUMNWISE else m)
    load().ora_damax(colrow, m, n, a.ctypes.data, _ld(a), v.ctypes.data)
    return v


def plant_mistake(which):
    load().ora_plant_mistake(which)


NOTRANS, TRANS = 111, 112
LEFT, RIGHT = 141, 142
NONUNIT, UNIT = 131, 132


def gemm(ta, tb, alpha, a, b, beta, c, nb, dtype=np.float64):
    """alpha op(a) op(b) + beta c by the tile algorithm (reference
    compute/pzgemm.c); returns the new c."""
    a, b = _f(a, dtype), _f(b, dtype)
    c = _f(c, dtype).copy(order="F")
    m, n = c.shape
    k = a.shape[1] if ta == NOTRANS else a.shape[0]
    fn = load().ora_dgemm if dtype == np.float64 else load().ora_sgemm
    rc = fn(ta, tb, m, n, k, alpha, a.ctypes.data, _ld(a), b.ctypes.data,
            _ld(b), beta, c.ctypes.data, _ld(c), nb)
    assert rc == 0, rc
    return c


def syrk(uplo, trans, alpha, a, beta, c, nb):
    """alpha a a^T + beta c (NOTRANS) or alpha a^T a + beta c (TRANS) in the
    uplo triangle of c (reference compute/pzsyrk.c); returns the new c."""
    a = _f(a, np.float64)
    c = _f(c, np.float64).copy(order="F")
    n = c.shape[0]
    k = a.shape[1] if trans == NOTRANS else a.shape[0]
    rc = load().ora_dsyrk(uplo, trans, n, k, alpha, a.ctypes.data, _ld(a),
                          beta, c.ctypes.data, _ld(c), nb)
    assert rc == 0, rc
    return c


def trsm(side, uplo, trans, diag, alpha, a, b, nb, dtype=np.float64):
    """alpha op(a)^-1 b (LEFT) or alpha b op(a)^-1 (RIGHT) by the tile
    algorithm (reference compute/pztrsm.c); returns the new b."""
    a = _f(a, dtype)
    b = _f(b, dtype).copy(order="F")
    m, n = b.shape
    fn = load().ora_dtrsm if dtype == np.float64 else load().ora_strsm
    rc = fn(side, uplo, trans, diag, m, n, alpha, a.ctypes.data, _ld(a),
            b.ctypes.data, _ld(b), nb)
    assert rc == 0, rc
    return b


# ---------------------------------------------------------------------------
# Complex precisions: dtype np.complex128 (z) or np.complex64 (c).

def _zc(dtype):
    return "z" if np.dtype(dtype) == np.complex128 else "c"


def zpotrf(uplo, a, nb, dtype=np.complex128):
    """Hermitian tile Cholesky (reference compute/pzpotrf.c); returns
    (factor array, info), the input is not modified."""
    a = _f(a, dtype).copy(order="F")
    info = getattr(load(), "ora_%spotrf" % _zc(dtype))(
        uplo, a.shape[0], a.ctypes.data, _ld(a), nb)
    return a, info


def zposv(uplo, a, b, nb, dtype=np.complex128):
    a = _f(a, dtype).copy(order="F")
    b = _f(b, dtype).copy(order="F")
    info = getattr(load(), "ora_%sposv" % _zc(dtype))(
        uplo, a.shape[0], b.shape[1], a.ctypes.data, _ld(a), b.ctypes.data,
        _ld(b), nb)
    return a, b, info


def zpotrs(uplo, l, b, nb, dtype=np.complex128):
    l = _f(l, dtype)
    b = _f(b, dtype).copy(order="F")
    info = getattr(load(), "ora_%spotrs" % _zc(dtype))(
        uplo, l.shape[0], b.shape[1], l.ctypes.data, _ld(l), b.ctypes.data,
        _ld(b), nb)
    return b, info


def zcposv(uplo, a, b, nb):
    """Mixed-precision Hermitian solve (reference compute/zcposv.c);
    returns (X, iter, info)."""
    a = _f(a, np.complex128)
    b = _f(b, np.complex128)
    x = np.zeros_like(b, order="F")
    it = ctypes.c_int(0)
    info = load().ora_zcposv(uplo, a.shape[0], b.shape[1], a.ctypes.data,
                             _ld(a), b.ctypes.data, _ld(b), x.ctypes.data,
                             _ld(x), ctypes.byref(it), nb)
    return x, it.value, info


def zlag2c(a):
    a = _f(a, np.complex128)
    s = np.zeros(a.shape, dtype=np.complex64, order="F")
    info = load().ora_zlag2c(a.shape[0], a.shape[1], a.ctypes.data, _ld(a),
                             s.ctypes.data, _ld(s))
    return s, info


def clag2z(s):
    s = _f(s, np.complex64)
    a = np.zeros(s.shape, dtype=np.complex128, order="F")
    load().ora_clag2z(s.shape[0], s.shape[1], s.ctypes.data, _ld(s),
                      a.ctypes.data, _ld(a))
    return a


def zlanhe_inf(uplo, a):
    a = _f(a, np.complex128)
    return load().ora_zlanhe_inf(uplo, a.shape[0], a.ctypes.data, _ld(a))


def dzamax(a):
    a = _f(a, np.complex128)
    v = np.zeros(a.shape[1])
    load().ora_dzamax(a.shape[0], a.shape[1], a.ctypes.data, _ld(a),
                      v.ctypes.data)
    return v
