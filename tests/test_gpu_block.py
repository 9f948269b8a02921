"""GPU: BLOCK mode (complex64 on FP32 MFMA, FP64 phase generation) vs the
oracle.  Bars: max |amplitude error| <= 2e-5 on random circuits (FP32 block
matrices and FP32 accumulation); the value-encoding circuit lands on its
basis state with fidelity >= 1 - 1e-6."""
import math

import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd.state import GpuState, Plan, QuantumState

from .helpers import random_ops
from .test_gpu_parity import run_gpu, run_oracle

pytestmark = pytest.mark.gpu
BLOCK = _lib.MODE_BLOCK


@pytest.mark.parametrize("n,seed", [(9, 1), (12, 2), (14, 3), (16, 4)])
def test_block_random_ops(gpu_lib, n, seed):
    rng = np.random.default_rng(seed)
    ops, ud = random_ops(rng, n, 80, with_blocks=False)
    st = W.normal_state(n, seed)
    re_o, im_o = run_oracle(n, ops, ud, st.re, st.im)
    re, im, stats = run_gpu(n, ops, ud, st.re.astype(np.float32), st.im.astype(np.float32), mode=BLOCK,
                            dtype=np.float32, renorm_every=-1)
    err = max(np.abs(re - re_o).max(), np.abs(im - im_o).max())
    assert err <= 2e-5, err


@pytest.mark.parametrize("n", [20, 24])
def test_block_value_encoding(gpu_lib, n):
    qc, v = W.value_encoding_precision(n, 77 + n)
    ops, ud = lower_circuit(qc)
    g = GpuState(n, np.float32)
    p = Plan(n, np.float32, ops, ud, _lib.opts(mode=BLOCK))
    p.execute(g)
    rev = int(format(v, f"0{n}b")[::-1], 2)
    pr = g.prob_mask((1 << n) - 1, rev)
    assert pr >= 1 - 1e-6, 1 - pr
    assert abs(g.norm2() - 1) < 1e-5
    # against the complex128 tile path on the same circuit
    ref = GpuState(n, np.float64)
    pf = Plan(n, np.float64, ops, ud, _lib.opts())
    pf.execute(ref)
    up = GpuState(n, np.float64)
    up.copy_from(g)
    l2 = math.sqrt(up.diff_norm2(ref))
    f = abs(up.inner(ref)) ** 2
    assert l2 < 2e-5 and f > 1 - 1e-6, (l2, f)
    for x in (g, ref, up):
        x.free()
    p.free()
    pf.free()


def test_block_renorm_and_layered(gpu_lib):
    n = 16
    ops, ud = lower_circuit(W.random_layered(n, 30, 9))
    z = np.zeros(1 << n)
    z[0] = 1
    re_o, im_o = run_oracle(n, ops, ud, z, np.zeros(1 << n))
    re, im, stats = run_gpu(n, ops, ud, z.astype(np.float32), np.zeros(1 << n, np.float32), mode=BLOCK,
                            dtype=np.float32)
    err = max(np.abs(re - re_o).max(), np.abs(im - im_o).max())
    assert err <= 2e-5, err
    assert abs(float(np.sum(re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2)) - 1) < 1e-6


@pytest.mark.parametrize("n,seed", [(9, 11), (11, 12), (14, 13), (17, 14)])
def test_block_complex128_random_ops(gpu_lib, n, seed):
    """complex128 blocks on FP64 MFMA: FP64 products and sums, only the
    summation order differs from the oracle's (<= 1e-12)."""
    rng = np.random.default_rng(seed)
    ops, ud = random_ops(rng, n, 90, with_blocks=False)
    st = W.normal_state(n, seed)
    re_o, im_o = run_oracle(n, ops, ud, st.re, st.im)
    re, im, _ = run_gpu(n, ops, ud, st.re.copy(), st.im.copy(), mode=BLOCK, dtype=np.float64)
    assert max(np.abs(re - re_o).max(), np.abs(im - im_o).max()) <= 1e-12


@pytest.mark.parametrize("n", [20, 24])
def test_block_complex128_value_encoding_and_renorm(gpu_lib, n):
    qc, v = W.value_encoding_precision(n, 5 + n)
    ops, ud = lower_circuit(qc)
    rev = int(format(v, f"0{n}b")[::-1], 2)
    for renorm in (0, 4):   # default (none) and a norm check every 4 blocks
        g = GpuState(n, np.float64)
        p = Plan(n, np.float64, ops, ud, _lib.opts(mode=BLOCK, renorm_every=renorm))
        p.execute(g)
        assert g.prob_mask((1 << n) - 1, rev) > 1 - 1e-12
        assert abs(g.norm2() - 1) < 1e-12
        p.free()
        g.free()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
@pytest.mark.parametrize("n,seed", [(10, 31), (15, 32)])
def test_block_dense_unitaries_on_mfma(gpu_lib, dtype, tol, n, seed):
    """Dense controlled unitaries folded into MFMA blocks vs the oracle."""
    rng = np.random.default_rng(seed)
    ops, ud = random_ops(rng, n, 90, with_blocks=True)
    st = W.normal_state(n, seed)
    re_o, im_o = run_oracle(n, ops, ud, st.re, st.im)
    re, im, _ = run_gpu(n, ops, ud, st.re.astype(dtype), st.im.astype(dtype), mode=BLOCK, dtype=dtype,
                        renorm_every=-1)
    assert max(np.abs(re - re_o).max(), np.abs(im - im_o).max()) <= tol
