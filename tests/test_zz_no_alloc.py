"""The allocation-free entry point hx3d_residuals_noalloc through the C ABI, against
hx3d_residuals (the entry point it is a variant of) and against max |QDot| of the derivative
read back: bit for bit (the maximum of doubles is exact), on two calls in a row, the second
on a smaller case, so that partial maxima a larger call left in the shared scratch array
cannot reach a later result."""
import ctypes as C

import numpy as np
import pytest

from horses3d_amd import mesh

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the engine library runs on (the same loaded instance): the test is the
    caller, and the caller owns its output buffer."""
    from horses3d_amd.engine import load_library
    load_library()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _noalloc(eng, out_dev):
    from horses3d_amd.engine import load_library
    rc = load_library().hx3d_residuals_noalloc(eng.h, out_dev)
    assert rc == 0, load_library().hx3d_last_error(eng.h)


@pytest.mark.parametrize("advance", [False, True])
def test_residuals_noalloc_matches_residuals(advance):
    from horses3d_amd.engine import Engine
    hip = _hip()
    out_dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(out_dev), 5 * 8) == 0   # caller-owned output
    kw = dict(mach=0.1, reynolds=1600.0, riemann="lxf", inviscid="split-form", averaging="pirozzoli")
    try:
        # a larger case first (2,000 blocks of 256 nodes), then a smaller one (27 elements, 7 blocks)
        for n, N in ((16, 4), (3, 3)):
            case = mesh.periodic_box(n, N, amplitude=0.1, phase_seed=2)
            print("case %d^3 x N=%d, advance=%s" % (n, N, advance), flush=True)
            eng = Engine(case, **kw)
            eng.set_state(mesh.taylor_green(case, 0.1, noise=1e-3, seed=n))
            if advance:
                eng.advance(2, 0.4, 0.4, t=0.0)   # the current QDot is the last stage's
            else:
                eng.qdot()
            _noalloc(eng, out_dev)
            eng.synchronize()
            got = np.full(5, -1.0)
            assert hip.hipMemcpy(got.ctypes.data, out_dev, 5 * 8, 2) == 0   # device to host
            ref = eng.residuals()
            qd = np.abs(eng.get_qdot()).max(axis=0)
            print("  noalloc", got, "residuals", ref, flush=True)
            assert np.array_equal(got, ref), (got, ref)
            assert np.array_equal(got, qd), (got, qd)
            eng.close()
    finally:
        hip.hipFree(out_dev)
