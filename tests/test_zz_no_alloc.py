"""charm_gpu_conservation_sums_noalloc against charm_gpu_conservation_sums,
through the C-ABI: two calls in a row on one GPU (the second on a smaller
mesh, with the other precision), each equal to the allocating entry point's
sums bit for bit.  The device output buffer comes from the same HIP runtime
the library links (libamdhip64), through ctypes."""
import ctypes as C

import numpy as np
import pytest

import charm_3d_amd as CA
import charm_3d_amd.workloads as W
from test_gpu_parity import device

pytestmark = pytest.mark.gpu

HIP_MEMCPY_D2H = 2


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def test_conservation_sums_noalloc_equals_allocating_entry_point():
    hip = _hip()
    L = CA.lib()
    fn = L.charm_gpu_conservation_sums_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p]
    for level, prec in ((6, CA.FP64), (3, CA.FP32)):
        print("[no_alloc] level %d precision %d" % (level, prec), flush=True)
        cfg, leaves, u0, per = W.c2(level=level, precision=prec)
        ctx = device(cfg, leaves, per, u0)
        ctx.run(ctx.get_dt(1), 2)
        ref = ctx.conservation_sums()
        nv = len(ref)
        dptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(dptr), 8 * nv) == 0
        try:
            CA.check(fn(ctx.h, dptr))
            ctx.synchronize()
            host = np.zeros(nv)
            assert hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), dptr, 8 * nv, HIP_MEMCPY_D2H) == 0
        finally:
            hip.hipFree(dptr)
        assert np.array_equal(host, ref), (host, ref)
        ctx.close()
