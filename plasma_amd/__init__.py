"""plasma_amd -- MI355X-native implementation of PLASMA's tile Cholesky
solve path (plasma_dpotrf / dpotrs / dposv / dsposv and the s twins).

The product is the C-ABI shared library ``plasma_amd/lib/libplasma_amd.so``
(public headers in ``include/``).  This module is a thin ctypes mirror of
that ABI for Python callers: the same function names, argument meaning and
return codes as the C interface (reference include/plasma_z.h:272-292,
include/plasma_zc.h:33-37).  It loads the HIP library and fails loudly when
it is missing; there is no CPU fallback.

Array arguments of the LAPACK-layout routines accept NumPy arrays (host
memory, staged through HBM) or torch tensors on the GPU (used in place).
Column-major storage is expected: pass ``np.asfortranarray(a)`` or, for
torch, a tensor whose element (i, j) sits at ``data_ptr + j*ld + i``
(e.g. ``a.t().contiguous()`` of a row-major matrix, shape (ncols, ld)).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libplasma_amd.so")

# --- enums (reference include/plasma_types.h:49-164) ---------------------
PlasmaRealFloat, PlasmaRealDouble = 2, 3
PlasmaComplexFloat, PlasmaComplexDouble = 4, 5
PlasmaNoTrans, PlasmaTrans, PlasmaConjTrans = 111, 112, 113
PlasmaUpper, PlasmaLower, PlasmaGeneral = 121, 122, 123
PlasmaGeneralBand = 124
PlasmaNonUnit, PlasmaUnit = 131, 132
PlasmaLeft, PlasmaRight = 141, 142
PlasmaInfNorm = 175
PlasmaColumnwise, PlasmaRowwise = 401, 402
PlasmaOneNorm, PlasmaFrobeniusNorm, PlasmaInfNorm, PlasmaMaxNorm = (171, 174,
                                                                    175, 177)
PlasmaSuccess = 0
PlasmaErrorNotInitialized = -101
PlasmaErrorNotSupported = -102
PlasmaErrorIllegalValue = -103
PlasmaErrorOutOfMemory = -104
PlasmaErrorInternal = -106
PlasmaErrorEnvironment = -109
PlasmaDisabled, PlasmaEnabled = 0, 1
PlasmaTuning, PlasmaNb, PlasmaIb = 0, 1, 2
PlasmaLookahead, PlasmaGraphCapture, PlasmaDevice = 1001, 1002, 1003
PlasmaUpdateDepth = 1004
PlasmaSplitUpdate = 1005
PlasmaHalfInput = 1006
PlasmaSplitBf16, PlasmaSplitFp16 = 2, 3

c_int, c_void_p, c_size_t, c_double = (ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_double)


class plasma_desc_t(ctypes.Structure):
    """Mirror of plasma_desc_t (include/plasma_descriptor.h)."""
    _fields_ = [("type", c_int), ("uplo", c_int), ("precision", c_int),
                ("matrix", c_void_p), ("A21", c_size_t), ("A12", c_size_t),
                ("A22", c_size_t), ("mb", c_int), ("nb", c_int),
                ("gm", c_int), ("gn", c_int), ("gmt", c_int), ("gnt", c_int),
                ("i", c_int), ("j", c_int), ("m", c_int), ("n", c_int),
                ("mt", c_int), ("nt", c_int), ("kl", c_int), ("ku", c_int),
                ("klt", c_int), ("kut", c_int)]


class plasma_request_t(ctypes.Structure):
    _fields_ = [("status", c_int)]


class plasma_sequence_t(ctypes.Structure):
    _fields_ = [("status", c_int), ("request", ctypes.POINTER(plasma_request_t))]


_D = plasma_desc_t
_PD = ctypes.POINTER(plasma_desc_t)
_PS = ctypes.POINTER(plasma_sequence_t)
_PR = ctypes.POINTER(plasma_request_t)
_PI = ctypes.POINTER(c_int)

# name -> (restype, argtypes); every entry point the headers declare
PROTOTYPES = {
    "plasma_init": (c_int, []),
    "plasma_finalize": (c_int, []),
    "plasma_set": (c_int, [c_int, c_int]),
    "plasma_get": (c_int, [c_int, _PI]),
    "plasma_version": (None, [_PI, _PI, _PI]),
    "plasma_uplo_const": (c_int, [ctypes.c_char]),
    "plasma_request_fail": (c_int, [_PS, _PR, c_int]),
    "plasma_request_init": (c_int, [_PR]),
    "plasma_sequence_init": (c_int, [_PS]),
    "plasma_sequence_sync": (c_int, [_PS]),
    "plasma_desc_general_create": (c_int, [c_int] * 9 + [_PD]),
    "plasma_desc_triangular_create": (c_int, [c_int] * 10 + [_PD]),
    "plasma_desc_destroy": (c_int, [_PD]),
    "plasma_desc_general_init": (c_int, [c_int, c_void_p] + [c_int] * 8 + [_PD]),
    "plasma_desc_triangular_init": (c_int, [c_int, c_int, c_void_p]
                                    + [c_int] * 8 + [_PD]),
    "plasma_desc_check": (c_int, [_D]),
    "plasma_desc_general_check": (c_int, [_D]),
    "plasma_desc_general_band_create": (c_int, [c_int] * 12 + [_PD]),
    "plasma_desc_general_band_init": (c_int, [c_int, c_int, c_void_p]
                                      + [c_int] * 10 + [_PD]),
    "plasma_desc_general_band_check": (c_int, [_D]),
    "plasma_desc_view": (_D, [_D, c_int, c_int, c_int, c_int]),
    "plasma_desc_storage_bytes": (c_size_t, [_D]),
}
for _p, _t in (("d", "double"), ("s", "float")):
    PROTOTYPES.update({
        "plasma_%spotrf" % _p: (c_int, [c_int, c_int, c_void_p, c_int]),
        "plasma_%spotrs" % _p: (c_int, [c_int, c_int, c_int, c_void_p, c_int,
                                        c_void_p, c_int]),
        "plasma_%sposv" % _p: (c_int, [c_int, c_int, c_int, c_void_p, c_int,
                                       c_void_p, c_int]),
        "plasma_omp_%spotrf" % _p: (None, [c_int, _D, _PS, _PR]),
        "plasma_omp_%spotrs" % _p: (None, [c_int, _D, _D, _PS, _PR]),
        "plasma_omp_%sposv" % _p: (None, [c_int, _D, _D, _PS, _PR]),
        "plasma_omp_%sge2desc" % _p: (None, [c_void_p, c_int, _D, _PS, _PR]),
        "plasma_omp_%sdesc2ge" % _p: (None, [_D, c_void_p, c_int, _PS, _PR]),
        "plasma_omp_%str2desc" % _p: (None, [c_void_p, c_int, _D, _PS, _PR]),
        "plasma_omp_%sdesc2tr" % _p: (None, [_D, c_void_p, c_int, _PS, _PR]),
        "plasma_%spotri" % _p: (c_int, [c_int, c_int, c_void_p, c_int]),
        "plasma_%spoinv" % _p: (c_int, [c_int, c_int, c_void_p, c_int]),
        "plasma_omp_%spotri" % _p: (None, [c_int, _D, _PS, _PR]),
        "plasma_omp_%spoinv" % _p: (None, [c_int, _D, _PS, _PR]),
        "plasma_%slansy" % _p: (c_double if _p == "d" else ctypes.c_float,
                                [c_int, c_int, c_int, c_void_p, c_int]),
        "plasma_%slange" % _p: (c_double if _p == "d" else ctypes.c_float,
                                [c_int, c_int, c_int, c_void_p, c_int]),
        "plasma_%samax" % _p: (c_int, [c_int, c_int, c_int, c_void_p, c_int,
                                       c_void_p]),
        "plasma_omp_%slansy" % _p: (None, [c_int, c_int, _D, c_void_p,
                                           c_void_p, _PS, _PR]),
        "plasma_omp_%slange" % _p: (None, [c_int, _D, c_void_p, c_void_p,
                                           _PS, _PR]),
        "plasma_omp_%samax" % _p: (None, [c_int, _D, c_void_p, c_void_p,
                                          _PS, _PR]),
    })
    _S = c_double if _p == "d" else ctypes.c_float
    PROTOTYPES.update({
        "plasma_%spbtrf" % _p: (c_int, [c_int, c_int, c_int, c_void_p, c_int]),
        "plasma_%spbtrs" % _p: (c_int, [c_int, c_int, c_int, c_int, c_void_p,
                                        c_int, c_void_p, c_int]),
        "plasma_%spbsv" % _p: (c_int, [c_int, c_int, c_int, c_int, c_void_p,
                                       c_int, c_void_p, c_int]),
        "plasma_omp_%spbtrf" % _p: (None, [c_int, _D, _PS, _PR]),
        "plasma_omp_%spbtrs" % _p: (None, [c_int, _D, _D, _PS, _PR]),
        "plasma_omp_%spbsv" % _p: (None, [c_int, _D, _D, _PS, _PR]),
        "plasma_omp_%spb2desc" % _p: (None, [c_void_p, c_int, _D, _PS, _PR]),
        "plasma_omp_%sdesc2pb" % _p: (None, [_D, c_void_p, c_int, _PS, _PR]),
    })
    PROTOTYPES.update({
        "plasma_%sgemm" % _p: (c_int, [c_int, c_int, c_int, c_int, c_int, _S,
                                       c_void_p, c_int, c_void_p, c_int, _S,
                                       c_void_p, c_int]),
        "plasma_omp_%sgemm" % _p: (None, [c_int, c_int, _S, _D, _D, _S, _D,
                                          _PS, _PR]),
        "plasma_%ssyrk" % _p: (c_int, [c_int, c_int, c_int, c_int, _S,
                                       c_void_p, c_int, _S, c_void_p, c_int]),
        "plasma_omp_%ssyrk" % _p: (None, [c_int, c_int, _S, _D, _S, _D, _PS,
                                          _PR]),
        "plasma_%strsm" % _p: (c_int, [c_int, c_int, c_int, c_int, c_int,
                                       c_int, _S, c_void_p, c_int, c_void_p,
                                       c_int]),
        "plasma_omp_%strsm" % _p: (None, [c_int, c_int, c_int, c_int, _S, _D,
                                          _D, _PS, _PR]),
    })
for _p in ("z", "c"):
    PROTOTYPES.update({
        "plasma_%spotrf" % _p: (c_int, [c_int, c_int, c_void_p, c_int]),
        "plasma_%spotrs" % _p: (c_int, [c_int, c_int, c_int, c_void_p, c_int,
                                        c_void_p, c_int]),
        "plasma_%sposv" % _p: (c_int, [c_int, c_int, c_int, c_void_p, c_int,
                                       c_void_p, c_int]),
        "plasma_omp_%spotrf" % _p: (None, [c_int, _D, _PS, _PR]),
        "plasma_omp_%spotrs" % _p: (None, [c_int, _D, _D, _PS, _PR]),
        "plasma_omp_%sposv" % _p: (None, [c_int, _D, _D, _PS, _PR]),
        "plasma_omp_%sge2desc" % _p: (None, [c_void_p, c_int, _D, _PS, _PR]),
        "plasma_omp_%sdesc2ge" % _p: (None, [_D, c_void_p, c_int, _PS, _PR]),
        "plasma_omp_%str2desc" % _p: (None, [c_void_p, c_int, _D, _PS, _PR]),
        "plasma_omp_%sdesc2tr" % _p: (None, [_D, c_void_p, c_int, _PS, _PR]),
    })
PROTOTYPES.update({
    "plasma_zcposv": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p,
                              c_int, c_void_p, c_int, _PI]),
    "plasma_omp_zcposv": (None, [c_int, _D, _D, _D, _D, _D, _D, c_void_p,
                                 c_void_p, c_void_p, _PI, _PS, _PR]),
    "plasma_zlag2c": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int]),
    "plasma_clag2z": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int]),
    "plasma_omp_zlag2c": (None, [_D, _D, _PS, _PR]),
    "plasma_omp_clag2z": (None, [_D, _D, _PS, _PR]),
})
PROTOTYPES["plasma_norm_work_doubles"] = (c_size_t, [_D])
PROTOTYPES.update({
    "plasma_dsposv": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p,
                              c_int, c_void_p, c_int, _PI]),
    "plasma_dhposv": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_void_p,
                              c_int, c_void_p, c_int, _PI]),
    "plasma_omp_dsposv": (None, [c_int, _D, _D, _D, _D, _D, _D, c_void_p,
                                 c_void_p, c_void_p, _PI, _PS, _PR]),
    "plasma_dlag2s": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int]),
    "plasma_slag2d": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int]),
    "plasma_omp_dlag2s": (None, [_D, _D, _PS, _PR]),
    "plasma_omp_slag2d": (None, [_D, _D, _PS, _PR]),
    "plasma_profile_enable": (c_int, [c_int]),
    "plasma_tuning_load": (c_int, [ctypes.c_char_p]),
    "plasma_dposv_vbatched": (c_int, [c_int, c_int, c_void_p, c_void_p,
                                      c_void_p, c_int, c_void_p, c_void_p,
                                      c_void_p]),
    "plasma_sposv_vbatched": (c_int, [c_int, c_int, c_void_p, c_void_p,
                                      c_void_p, c_int, c_void_p, c_void_p,
                                      c_void_p]),
    "plasma_dposv_vbatched_noalloc": (c_int, [c_int, c_int, c_void_p,
                                              c_void_p, c_void_p, c_int,
                                              c_void_p, c_void_p, c_void_p]),
    "plasma_dpotrf_vbatched": (c_int, [c_int, c_int, c_void_p, c_void_p,
                                       c_void_p, c_void_p]),
    "plasma_spotrf_vbatched": (c_int, [c_int, c_int, c_void_p, c_void_p,
                                       c_void_p, c_void_p]),
    "plasma_dpotrf_adjoint": (c_int, [c_int, c_int, c_void_p, c_int,
                                      c_void_p, c_int]),
    "plasma_spotrf_adjoint": (c_int, [c_int, c_int, c_void_p, c_int,
                                      c_void_p, c_int]),
    "plasma_dpotrs_adjoint": (c_int, [c_int, c_int, c_int, c_void_p, c_int,
                                      c_void_p, c_int, c_void_p, c_int,
                                      c_void_p, c_int]),
    "plasma_spotrs_adjoint": (c_int, [c_int, c_int, c_int, c_void_p, c_int,
                                      c_void_p, c_int, c_void_p, c_int,
                                      c_void_p, c_int]),
    "plasma_profile_query": (c_int, [c_int, ctypes.POINTER(c_double),
                                     ctypes.POINTER(c_double),
                                     ctypes.POINTER(c_double),
                                     ctypes.POINTER(ctypes.c_long)]),
})

_lib = None


def load():
    """Load libplasma_amd.so and attach prototypes (raises if missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libplasma_amd.so is not built: run "
                           "`python -m plasma_amd.build` (needs hipcc)")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(x):
    """Device or host address of a torch tensor / NumPy array / int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError("unsupported array type %r" % type(x))


def init():
    return load().plasma_init()


def finalize():
    return load().plasma_finalize()


def set_param(param, value):
    return load().plasma_set(param, value)


def get_param(param):
    v = c_int(0)
    rc = load().plasma_get(param, ctypes.byref(v))
    return rc, v.value


def potrf(uplo, n, a, lda, prec="d"):
    return getattr(load(), "plasma_%spotrf" % prec)(uplo, n, _ptr(a), lda)


def potrs(uplo, n, nrhs, a, lda, b, ldb, prec="d"):
    return getattr(load(), "plasma_%spotrs" % prec)(uplo, n, nrhs, _ptr(a),
                                                     lda, _ptr(b), ldb)


def posv(uplo, n, nrhs, a, lda, b, ldb, prec="d"):
    return getattr(load(), "plasma_%sposv" % prec)(uplo, n, nrhs, _ptr(a), lda,
                                                    _ptr(b), ldb)


def potrf_adjoint(uplo, n, a, lda, abar, ldabar, prec="d"):
    return getattr(load(), "plasma_%spotrf_adjoint" % prec)(
        uplo, n, _ptr(a), lda, _ptr(abar), ldabar)


def potrs_adjoint(uplo, n, nrhs, a, lda, x, ldx, xbar, ldxbar, abar, ldabar,
                  prec="d"):
    return getattr(load(), "plasma_%spotrs_adjoint" % prec)(
        uplo, n, nrhs, _ptr(a), lda, _ptr(x), ldx, _ptr(xbar), ldxbar,
        _ptr(abar), ldabar)


def dsposv(uplo, n, nrhs, a, lda, b, ldb, x, ldx):
    it = c_int(0)
    rc = load().plasma_dsposv(uplo, n, nrhs, _ptr(a), lda, _ptr(b), ldb,
                              _ptr(x), ldx, ctypes.byref(it))
    return rc, it.value


def zcposv(uplo, n, nrhs, a, lda, b, ldb, x, ldx):
    """plasma_zcposv: (return code, iterations)."""
    it = c_int(0)
    rc = load().plasma_zcposv(uplo, n, nrhs, _ptr(a), lda, _ptr(b), ldb,
                              _ptr(x), ldx, ctypes.byref(it))
    return rc, it.value


def dhposv(uplo, n, nrhs, a, lda, b, ldb, x, ldx):
    """plasma_dhposv: (return code, total inner GMRES iterations or -1)."""
    it = c_int(0)
    rc = load().plasma_dhposv(uplo, n, nrhs, _ptr(a), lda, _ptr(b), ldb,
                              _ptr(x), ldx, ctypes.byref(it))
    return rc, it.value
