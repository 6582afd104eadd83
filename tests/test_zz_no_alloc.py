"""fvg_conserved_totals_noalloc (no reference counterpart: the allocation-free variant of fvg_conserved_totals for
callers that cannot allocate): the same bits as the context entry point on two calls in a row, the second on a
smaller state, with the state and the output in device memory the caller owns."""
import ctypes as C

import numpy as np
import pytest

from cases import base_keys, smooth_state
from fenice_amd import fvg

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself runs on (already loaded with libfvg.so), for device memory and a stream
    that the test owns as a caller would."""
    fvg.lib()
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    H.hipStreamSynchronize.argtypes = [C.c_void_p]
    H.hipStreamDestroy.argtypes = [C.c_void_p]
    H.hipSetDevice.argtypes = [C.c_int]
    return H


H2D, D2H = 1, 2


def _context_totals(keys, U):
    with fvg.Solver(fvg.config_from_keys(keys), device=0) as s:
        s.upload(U)
        return s.conserved_totals()


def _noalloc_totals(H, U, ncomp, stream):
    planes = np.ascontiguousarray(U[:, :ncomp].T)
    out = np.full(3, np.nan)
    dU, dO = C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(dU), planes.nbytes) == 0 and H.hipMalloc(C.byref(dO), out.nbytes) == 0
    try:
        assert H.hipMemcpy(dU, planes.ctypes.data, planes.nbytes, H2D) == 0
        assert H.hipMemcpy(dO, out.ctypes.data, out.nbytes, H2D) == 0
        rc = fvg.lib().fvg_conserved_totals_noalloc(0, dU, U.shape[0], ncomp, dO, stream)
        assert rc == fvg.OK, rc
        assert H.hipStreamSynchronize(stream) == 0
        assert H.hipMemcpy(out.ctypes.data, dO, out.nbytes, D2H) == 0
    finally:
        H.hipFree(dU)
        H.hipFree(dO)
    return out


def test_noalloc_totals_equal_the_context_entry_point():
    H = _hip()
    assert H.hipSetDevice(0) == 0
    stream = C.c_void_p()
    assert H.hipStreamCreate(C.byref(stream)) == 0
    cases = [(base_keys(40, 36, 32), smooth_state(40, 36, 32, 3), 5),   # 3-D, then a smaller 2-D state
             (base_keys(24, 20), smooth_state(24, 20, 1, 4), 4)]
    try:
        for keys, U, ncomp in cases:
            print("noalloc totals", keys["nx"], keys["ny"], keys["nz"], flush=True)
            want = _context_totals(keys, U)
            got = _noalloc_totals(H, U, ncomp, stream)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    finally:
        H.hipStreamDestroy(stream)


def test_noalloc_totals_refuse_bad_sizes():
    H = _hip()
    x, o = C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(x), 40) == 0 and H.hipMalloc(C.byref(o), 24) == 0
    L = fvg.lib()
    try:
        assert L.fvg_conserved_totals_noalloc(0, x, 0, 5, o, None) == fvg.EINVAL
        assert L.fvg_conserved_totals_noalloc(0, x, 1 << 31, 5, o, None) == fvg.EINVAL
        assert L.fvg_conserved_totals_noalloc(0, x, 1, 3, o, None) == fvg.EINVAL
        assert L.fvg_conserved_totals_noalloc(0, None, 1, 5, o, None) == fvg.EINVAL
        assert L.fvg_conserved_totals_noalloc(-1, x, 1, 5, o, None) == fvg.EINVAL
    finally:
        H.hipFree(x)
        H.hipFree(o)
