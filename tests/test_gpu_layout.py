"""GPU tests of the layout translation kernels (bit-exact).

The tile storage image written by plasma_omp_d{tr2desc,ge2desc} on the
device must equal, element for element, the image the oracle builds with
the reference's tile addressing (include/plasma_descriptor.h:93-151,
compute/pztr2desc.c, pzge2desc.c), and desc2tr / desc2ge must restore the
LAPACK array, leaving entries outside the translated region untouched.
"""
import ctypes

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu


def _seq():
    s, r = pa.plasma_sequence_t(), pa.plasma_request_t()
    lib = pa.load()
    lib.plasma_sequence_init(ctypes.byref(s))
    lib.plasma_request_init(ctypes.byref(r))
    return s, r


@pytest.mark.parametrize("region", [pa.PlasmaLower, pa.PlasmaUpper,
                                    pa.PlasmaGeneral])
@pytest.mark.parametrize("m,n,nb", [(100, 100, 16), (257, 257, 64),
                                    (1000, 1000, 256), (300, 70, 128)])
@pytest.mark.parametrize("device_ptr", [False, True])
def test_layout_roundtrip(plasma, region, m, n, nb, device_ptr):
    if region != pa.PlasmaGeneral and m != n:
        pytest.skip("triangular descriptors are square")
    lib = pa.load()
    rng = np.random.default_rng(m * 7 + n)
    lda = m + 5
    A = np.asfortranarray(rng.standard_normal((lda, n)))
    desc = pa.plasma_desc_t()
    if region == pa.PlasmaGeneral:
        rc = lib.plasma_desc_general_create(pa.PlasmaRealDouble, nb, nb, m, n,
                                            0, 0, m, n, ctypes.byref(desc))
    else:
        rc = lib.plasma_desc_triangular_create(pa.PlasmaRealDouble, region,
                                               nb, nb, m, n, 0, 0, m, n,
                                               ctypes.byref(desc))
    assert rc == 0
    try:
        s, r = _seq()
        if device_ptr:
            import torch
            dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
            torch.cuda.synchronize()
            src = dA.data_ptr()
        else:
            src = A.ctypes.data
        if region == pa.PlasmaGeneral:
            lib.plasma_omp_dge2desc(src, lda, desc, ctypes.byref(s),
                                    ctypes.byref(r))
        else:
            lib.plasma_omp_dtr2desc(src, lda, desc, ctypes.byref(s),
                                    ctypes.byref(r))
        assert lib.plasma_sequence_sync(ctypes.byref(s)) == 0
        count = lib.plasma_desc_storage_bytes(desc) // 8
        img = U.d2h(desc.matrix, count)
        oregion = {pa.PlasmaLower: ora.LOWER, pa.PlasmaUpper: ora.UPPER,
                   pa.PlasmaGeneral: 0}[region]
        ref = ora.lapack2tile(oregion, A[:m, :], nb)
        assert ref.shape[0] == count
        if region == pa.PlasmaGeneral:
            assert np.array_equal(img, ref)
        else:
            # compare the stored tiles (edge-region padding is unused)
            back = ora.tile2lapack(oregion, img, m, n, nb, init=np.zeros((m, n)))
            want = ora.tile2lapack(oregion, ref, m, n, nb, init=np.zeros((m, n)))
            assert np.array_equal(back, want)
        # translate back into a fresh array: region entries restored,
        # everything else untouched
        out = np.asfortranarray(np.full((lda, n), -7.0))
        s, r = _seq()
        if region == pa.PlasmaGeneral:
            lib.plasma_omp_ddesc2ge(desc, out.ctypes.data, lda,
                                    ctypes.byref(s), ctypes.byref(r))
        else:
            lib.plasma_omp_ddesc2tr(desc, out.ctypes.data, lda,
                                    ctypes.byref(s), ctypes.byref(r))
        assert lib.plasma_sequence_sync(ctypes.byref(s)) == 0
        want = ora.tile2lapack(oregion, ref, m, n, nb,
                               init=np.full((m, n), -7.0))
        assert np.array_equal(out[:m, :], want)
        assert np.all(out[m:, :] == -7.0)
    finally:
        lib.plasma_desc_destroy(ctypes.byref(desc))
