"""GPU: bfly_prob_mask_noalloc, the allocation-free entry point, against
bfly_prob_mask through the C ABI.

Both run the same reduction kernels with the same fixed block count per size,
so the results must be bit-identical (the parity bar for this integer-ordered
FP64 reduction).  Two calls in a row, the second on a smaller state (the
static scratch array keeps the first call's partials: the second must not
read them), and in both precisions.  The result goes to a device address the
test allocates (the entry point itself allocates nothing)."""
import ctypes as ct

import numpy as np
import pytest

from code_mojo_amd import _lib
from code_mojo_amd.state import GpuState, dtype_code

pytestmark = pytest.mark.gpu


def _hip():
    hip = ct.CDLL("libamdhip64.so.7")   # the runtime libbfly_hip.so is linked against
    hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    hip.hipFree.argtypes = [ct.c_void_p]
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = ct.c_int
    return hip


def noalloc(lib, hip, g, mask, value, d_out):
    pr, pi = ct.c_void_p(), ct.c_void_p()
    _lib.check(lib.bfly_state_device_ptrs(g.handle, ct.byref(pr), ct.byref(pi)))
    g.synchronize()                     # the state's own stream is idle
    n = g.n
    _lib.check(lib.bfly_prob_mask_noalloc(pr, pi, n, dtype_code(g.dtype), mask, value, d_out, None))
    assert hip.hipDeviceSynchronize() == 0
    out = ct.c_double()
    assert hip.hipMemcpy(ct.byref(out), d_out, 8, 2) == 0
    return out.value


def test_noalloc_equals_prob_mask_on_consecutive_calls(gpu_lib):
    lib = _lib.load()
    hip = _hip()
    d_out = ct.c_void_p()
    assert hip.hipMalloc(ct.byref(d_out), 64) == 0
    try:
        for n, dtype, seed in [(22, np.float64, 1), (13, np.float32, 2)]:   # the second call is smaller
            print(f"[noalloc] n={n} dtype={np.dtype(dtype).name}", flush=True)
            g = GpuState(n, dtype)
            g.fill_random(seed)
            cases = [(0, 0), ((1 << n) - 1, 5), (0b1011, 0b0010), (1 << (n - 1), 1 << (n - 1))]
            for mask, value in cases:
                got = noalloc(lib, hip, g, mask, value, d_out)
                ref = g.prob_mask(mask, value)
                assert got == ref, (n, mask, value, got, ref)
            # a state normalised in its own precision (float32 rounding for complex64)
            assert abs(noalloc(lib, hip, g, 0, 0, d_out) - 1.0) < (1e-12 if dtype == np.float64 else 1e-6)
            g.free()
        # refused inputs: nothing is launched
        g = GpuState(10)
        pr, pi = ct.c_void_p(), ct.c_void_p()
        _lib.check(lib.bfly_state_device_ptrs(g.handle, ct.byref(pr), ct.byref(pi)))
        assert lib.bfly_prob_mask_noalloc(pr, pi, 41, 0, 0, 0, d_out, None) == _lib.BFLY_EINVAL
        assert lib.bfly_prob_mask_noalloc(pr, pi, 10, 0, 0b01, 0b10, d_out, None) == _lib.BFLY_EINVAL
        assert lib.bfly_prob_mask_noalloc(pr, pi, 10, 7, 0, 0, d_out, None) == _lib.BFLY_EINVAL
        assert lib.bfly_prob_mask_noalloc(pr, pi, 10, 0, 0, 0, None, None) == _lib.BFLY_EINVAL
        g.free()
    finally:
        hip.hipFree(d_out)
