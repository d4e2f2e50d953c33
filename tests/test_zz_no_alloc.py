"""pqdts_simplexproj_rows_noalloc against pqdts_simplexproj_rows, through the C ABI.

Device buffers come from the HIP runtime that libpqdts_amd.so itself links
(libamdhip64.so.7, opened by ctypes), so that the pointers belong to the same
runtime. Two calls in a row, the second smaller and in place; the rows must be
bit-identical to the allocating entry point's. A call beyond the row limit fails
with an error."""
import ctypes as C

import numpy as np
import pytest

import pqdts_amd
from pqdts_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.load()  # loads libamdhip64.so.7 as a dependency first
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _dev(h, nbytes):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), nbytes) == 0
    return p


def test_noalloc_projection_matches_allocating_entry_point():
    h = _hip()
    L = pqdts_amd.load()
    rng = np.random.default_rng(31)
    cap = 3000 * 257
    y_d, x_d = _dev(h, 8 * cap), _dev(h, 8 * cap)
    try:
        for rows, n, in_place in ((3000, 257, False), (700, 64, True)):
            Y = rng.normal(size=(rows, n)) * 0.3 + 1.0 / n
            Y[0] = -1.0          # a row of equal negatives
            Y[1] = 1.0 / n       # a row already on the simplex
            ref = pqdts_amd.simplexproj_rows(Y)
            assert h.hipMemcpy(y_d, Y.ctypes.data, Y.nbytes, H2D) == 0
            out_d = y_d if in_place else x_d
            rc = L.pqdts_simplexproj_rows_noalloc(y_d, out_d, rows, n, None)
            assert rc == 0, L.pqdts_last_error()
            assert h.hipDeviceSynchronize() == 0
            got = np.empty_like(Y)
            assert h.hipMemcpy(got.ctypes.data, out_d, got.nbytes, D2H) == 0
            assert np.array_equal(got, ref), (rows, n, float(np.max(np.abs(got - ref))))
            if not in_place:
                back = np.empty_like(Y)
                assert h.hipMemcpy(back.ctypes.data, y_d, back.nbytes, D2H) == 0
                assert np.array_equal(back, Y)  # y untouched
        rc = L.pqdts_simplexproj_rows_noalloc(y_d, x_d, (1 << 20) + 1, 1, None)
        assert rc != 0 and b"1048576" in L.pqdts_last_error()
    finally:
        h.hipFree(y_d)
        h.hipFree(x_d)
