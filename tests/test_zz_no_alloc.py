"""GPU: hbi_hmat_matvec_noalloc, the product that allocates nothing, against
hbi_hmat_matvec through the C ABI.

Two calls in a row share the library's scratch array, the second on a
smaller matrix; FP64, FP32 and mixed factors.  Both entry points run the same
kernels on the same plan, so the results are bit-identical (integer-exact
comparison).  Device buffers are the test's own, from the HIP runtime."""
import ctypes as C

import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(8 * n)) == 0
        return p

    def put(self, p, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def get(self, p, n):
        a = np.empty(n)
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0
        return a


def test_gpu_matvec_noalloc_matches_matvec():
    hip = _Hip()
    L = H.lib()
    rng = np.random.default_rng(11)
    cases = [("3dph 96x48", P.mesh_3dph(96, 48, 1.0), 32, 1e-5), ("2dvs 1024", P.mesh_2dvs(1024, 40.0 / 1024), 32, 1e-6)]
    for label, g, leaf, eps in cases:
        Hg = H.HMatrix(g, leaf=leaf, eps=eps, bf16_margin=1.0)
        n = g.n
        x = rng.standard_normal(n)
        dx, dy1, dy2 = hip.alloc(n), hip.alloc(n), hip.alloc(n)
        try:
            hip.put(dx, x)
            for prec in (H.PREC_F64, H.PREC_F32, H.PREC_MIXED):
                print("%s, %d elements, precision %d" % (label, n, prec), flush=True)
                y_host = Hg.matvec(x, prec)              # builds the precision's layout
                dp = C.POINTER(C.c_double)
                assert L.hbi_hmat_matvec(Hg.h, C.cast(dx, dp), C.cast(dy1, dp), prec, 1) == 0, L.hbi_last_error()
                hip.put(dy2, np.full(n, np.nan))
                assert L.hbi_hmat_matvec_noalloc(Hg.h, dx, dy2, prec) == 0, L.hbi_last_error()
                y1, y2 = hip.get(dy1, n), hip.get(dy2, n)
                assert np.array_equal(y1, y2) and np.array_equal(y1, y_host)
        finally:
            for p in (dx, dy1, dy2):
                hip.L.hipFree(p)
        Hg.close()


def test_gpu_matvec_noalloc_rejects_bad_arguments():
    """A bad precision or a NULL vector is refused before any launch; the
    product still works afterwards."""
    hip = _Hip()
    L = H.lib()
    g = P.mesh_2dvs(512, 40.0 / 512)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    dx, dy = hip.alloc(g.n), hip.alloc(g.n)
    try:
        x = np.linspace(-1.0, 1.0, g.n)
        hip.put(dx, x)
        assert L.hbi_hmat_matvec_noalloc(Hg.h, dx, dy, 7) == H.E_ARG
        assert b"precision" in L.hbi_last_error()
        assert L.hbi_hmat_matvec_noalloc(Hg.h, None, dy, H.PREC_F64) == H.E_ARG
        assert L.hbi_hmat_matvec_noalloc(None, dx, dy, H.PREC_F64) == H.E_ARG
        y = Hg.matvec(x, H.PREC_MIXED)
        assert L.hbi_hmat_matvec_noalloc(Hg.h, dx, dy, H.PREC_MIXED) == 0
        assert np.array_equal(hip.get(dy, g.n), y)
    finally:
        hip.L.hipFree(dx)
        hip.L.hipFree(dy)
    Hg.close()
