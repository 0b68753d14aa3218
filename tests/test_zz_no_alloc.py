"""clg_multidot_noalloc, the multi-dot product that allocates nothing (its partial sums live in a static
device array), against clg_multidot on the same fields: two calls in a row, the second smaller, in FP64
and FP32.  Both run the same kernels on the same grid, so the sums agree to the last bit; the bar
written here is 1e-15 relative."""
import ctypes as C

import numpy as np
import pytest

import cudalatticegauge_amd as clg

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libclg_hip.so itself uses (already loaded by it): the caller's own device buffer
    for the result comes from there, so no second runtime enters the process."""
    clg.lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _noalloc_multidot(ctx, V, w):
    hip = _hip()
    nb = 16 * len(V)
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), nb) == 0
    try:
        hs = (C.c_void_p * len(V))(*[v.h for v in V])
        clg.check(clg.lib().clg_multidot_noalloc(ctx.h, len(V), hs, w.h, dev))
        ctx.sync()
        host = (C.c_double * (2 * len(V)))()
        assert hip.hipMemcpy(host, dev, nb, 2) == 0  # device to host
    finally:
        hip.hipFree(dev)
    o = np.array(host[:])
    return o[0::2] + 1j * o[1::2]


@pytest.mark.parametrize("prec", [clg.F64, clg.F32])
def test_noalloc_multidot_matches_multidot(prec):
    with clg.Context(0) as ctx:
        # first call: 20 vectors (two 16-vector chunks) on 8^4; second: 3 vectors on 4^4
        for dims, nv in (((8, 8, 8, 8), 20), ((4, 4, 4, 4), 3)):
            print(f"noalloc multidot prec {prec} dims {dims} n {nv}", flush=True)
            V = []
            for k in range(nv):
                f = clg.Field(ctx, clg.WILSON, prec, dims)
                f.gaussian(100 + k)
                V.append(f)
            w = clg.Field(ctx, clg.WILSON, prec, dims)
            w.gaussian(7)
            ref = clg.multidot(ctx, V, w)
            got = _noalloc_multidot(ctx, V, w)
            assert np.all(np.isfinite(got))
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err <= 1e-15, f"dims {dims} n {nv}: relative difference {err:.3e}"


def test_noalloc_multidot_rejects_sizes_beyond_the_static_buffer():
    with clg.Context(0) as ctx:
        dims = (4, 4, 4, 4)
        w = clg.Field(ctx, clg.WILSON, clg.F64, dims)
        hs = (C.c_void_p * 65)(*([w.h] * 65))
        assert clg.lib().clg_multidot_noalloc(ctx.h, 65, hs, w.h, C.c_void_p(1)) == 2
        assert b"n must be in [1, 64]" in clg.lib().clg_last_error()
