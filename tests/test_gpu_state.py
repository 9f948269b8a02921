"""GPU tests of state handles: bfly_state_copy is ordered against the
source's later work (the copy reads the source as it was at the call), and a
copy onto itself is refused."""
import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd._lib import BflyError
from code_mojo_amd.state import GpuState, Plan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dst_dtype", [np.float64, np.float32])
def test_copy_then_plan_on_source_keeps_the_copy(gpu_lib, dst_dtype):
    n = 24                      # 256 MiB of complex128: the copy takes long enough to race
    src = GpuState(n)
    src.upload(W.normal_state(n, 5))
    before = src.download()
    dst = GpuState(n, dst_dtype)
    ops, ud = lower_circuit(W.random_layered(n, 4, 9))
    plan = Plan(n, np.float64, ops, ud, _lib.opts())
    dst.copy_from(src)          # asynchronous on dst's stream
    plan.execute(src)           # at once, on src's stream
    got = dst.download()
    after = src.download()
    assert not np.array_equal(after.re, before.re)          # the plan did change src
    if dst_dtype == np.float64:
        assert np.array_equal(got.re, before.re) and np.array_equal(got.im, before.im)
    else:
        assert np.array_equal(got.re, before.re.astype(np.float32))
        assert np.array_equal(got.im, before.im.astype(np.float32))


def test_copy_waits_for_pending_source_work(gpu_lib):
    n = 22
    src = GpuState(n)
    src.upload(W.normal_state(n, 6))
    ops, ud = lower_circuit(W.random_layered(n, 3, 1))
    plan = Plan(n, np.float64, ops, ud, _lib.opts())
    dst = GpuState(n)
    plan.execute(src)           # queued on src's stream
    dst.copy_from(src)          # must see the plan's result
    assert np.array_equal(dst.download().re, src.download().re)


def test_copy_onto_itself_is_refused(gpu_lib):
    g = GpuState(8)
    with pytest.raises(BflyError):
        g.copy_from(g)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 2, 3, 12, 21])
def test_scale_streaming_kernel(gpu_lib, dtype, n):
    # 16-byte vectors plus the scalar tail of states smaller than one vector
    g = GpuState(n, dtype)
    st = W.normal_state(n, 40 + n, dtype)
    g.upload(st)
    g.scale(-0.375)             # exact in both precisions: results must equal the host product
    out = g.download()
    g.free()
    assert np.array_equal(out.re, (st.re * dtype(-0.375)).astype(dtype))
    assert np.array_equal(out.im, (st.im * dtype(-0.375)).astype(dtype))
