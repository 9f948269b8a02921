"""GPU: phase flips, diffusers and the Grover peephole vs the gate-by-gate
oracle.  complex128 <= 1e-12 max |amplitude error|; complex64 <= 2e-5 on
Grover states (a few thousand FP32 roundings of O(2^-n/2) amplitudes)."""
import math

import numpy as np
import pytest

from code_mojo_amd import ClassicalReplacementKind as CRK
from code_mojo_amd import _lib, algos, lower_circuit, replace_quantum_with_classical
from code_mojo_amd import workloads as W
from code_mojo_amd.lowering import OpList

from .helpers import random_ops
from .test_gpu_parity import TOL_C128, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,dtype,tol", [(14, np.float64, TOL_C128), (16, np.float64, TOL_C128),
                                         (15, np.float32, 2e-5), (17, np.float32, 2e-5)])
def test_grover_peephole_vs_oracle(gpu_lib, n, dtype, tol):
    qc, marked = W.grover_search(n, 40 + n)
    ops, ud = lower_circuit(qc)
    z = np.zeros(1 << n)
    z[0] = 1
    re, im, st = run_gpu(n, ops, ud, z.copy(), np.zeros(1 << n), dtype=dtype)
    re_o, im_o = run_oracle(n, ops, ud, z, np.zeros(1 << n))
    err = max(np.abs(re - re_o).max(), np.abs(im - im_o).max())
    assert err <= tol, err
    assert re[marked] ** 2 + im[marked] ** 2 > 0.99
    iters = int(math.floor(math.pi / 4 * math.sqrt(2 ** n)))
    assert st["launches"] < 3 * iters  # oracle + diffuser per iteration, not tile passes


def test_classical_diffuser_circuit(gpu_lib):
    n = 14
    marked = 12345
    qc = algos.grover_search_circuit(n, marked, 20)
    replace_quantum_with_classical(qc, [CRK.GROVER_DIFFUSER])
    ops, ud = lower_circuit(qc)
    st = W.normal_state(n, 3)
    re, im, _ = run_gpu(n, ops, ud, st.re.copy(), st.im.copy())
    re_o, im_o = run_oracle(n, ops, ud, st.re, st.im)
    # the reference's classical diffuser maps im to -2 mean_im - im, which is
    # not norm-preserving on complex states (here |psi| grows to ~1e7), so the
    # bar is relative to the largest amplitude
    scale = max(np.abs(re_o).max(), np.abs(im_o).max(), 1.0)
    assert max(np.abs(re - re_o).max(), np.abs(im - im_o).max()) <= TOL_C128 * scale


@pytest.mark.parametrize("mask", [0b1111111111111111, 0b0000111111111111, 0b1010101010101010, 0b0000000000000110,
                                  0b1000000000000001])
@pytest.mark.parametrize("variant", [0, 1])
def test_diffuse_masks_on_complex_states(gpu_lib, mask, variant):
    n = 16
    out = OpList(n)
    out.diffuse(mask, variant)
    out.phase_flip(mask, mask & 0x5A5A)
    out.diffuse(mask, variant)
    ops, ud = out.finish()
    st = W.normal_state(n, mask % 97)
    for dtype, tol in ((np.float64, TOL_C128), (np.float32, 2e-6)):
        re, im, _ = run_gpu(n, ops, ud, st.re.copy(), st.im.copy(), dtype=dtype)
        re_o, im_o = run_oracle(n, ops, ud, st.re, st.im)
        assert max(np.abs(re - re_o).max(), np.abs(im - im_o).max()) <= tol, dtype


def _spread(n, t, seed):
    rng = np.random.default_rng(seed)
    bits = rng.choice(n, size=t, replace=False)
    return int(sum(1 << int(b) for b in bits))


# partial diffusers over more than 12 qubits (apply_grover_diffuser takes any
# target list): few blocks split into chunks (two passes), and 2^14 or more
# blocks of 2^t amplitudes (one workgroup per block)
@pytest.mark.parametrize("n,mask", [(16, (1 << 14) - 1), (20, _spread(20, 15, 1)), (22, ((1 << 21) - 1) << 1),
                                    (18, ((1 << 13) - 1) << 5), (27, _spread(27, 13, 2)), (27, (1 << 13) - 1)])
@pytest.mark.parametrize("variant", [0, 1])
def test_wide_partial_diffusers(gpu_lib, n, mask, variant):
    out = OpList(n)
    out.diffuse(mask, variant)
    out.phase_flip(mask, mask & 0x2A5A5A5)
    out.diffuse(mask, variant)
    ops, ud = out.finish()
    print(f"[wide-diffuse] n={n} mask={mask:#x} variant={variant}", flush=True)
    # a random state plus a common offset: the block means are of the order
    # of the amplitudes, so a wrong or missing mean shows at every tolerance
    st = W.normal_state(n, n + variant, np.float32 if n > 24 else np.float64)
    off = 2.0 ** (-n / 2)
    re0 = (st.re.astype(np.float64) + 3 * off).astype(st.re.dtype)
    im0 = (st.im.astype(np.float64) - off).astype(st.im.dtype)
    re_o, im_o = run_oracle(n, ops, ud, re0.astype(np.float64), im0.astype(np.float64))
    scale = max(np.abs(re_o).max(), np.abs(im_o).max())
    # complex128: the oracle bar; complex64: 1e-5 of the largest amplitude
    dtypes = ((np.float32, 1e-5 * scale),) if n > 24 else ((np.float64, TOL_C128), (np.float32, 1e-5 * scale))
    for dtype, tol in dtypes:
        re, im, _ = run_gpu(n, ops, ud, re0.astype(dtype), im0.astype(dtype), dtype=dtype)
        err = max(np.abs(re - re_o).max(), np.abs(im - im_o).max())
        assert err <= tol, (dtype, err, tol)


@pytest.mark.parametrize("seed", [5, 6])
def test_flips_inside_tile_passes(gpu_lib, seed):
    n = 15
    rng = np.random.default_rng(seed)
    ops, ud = random_ops(rng, n, 80, with_blocks=False, with_perms=False)
    extra = OpList(n)
    for _ in range(10):
        mask = int(rng.integers(1, 1 << n))
        extra.phase_flip(mask, int(rng.integers(0, 1 << n)) & mask)
    eops, _ = extra.finish()
    allops = np.concatenate([ops[:30], eops[:5], ops[30:60], eops[5:], ops[60:]])
    st = W.normal_state(n, seed)
    re_o, im_o = run_oracle(n, allops, ud, st.re, st.im)
    for tb in (0, 9):
        re, im, _ = run_gpu(n, allops, ud, st.re.copy(), st.im.copy(), tile_bits=tb)
        assert max(np.abs(re - re_o).max(), np.abs(im - im_o).max()) <= TOL_C128, tb


def test_diffuser_deterministic(gpu_lib):
    n = 18
    qc, _ = W.grover_search(n, 9)
    ops, ud = lower_circuit(qc)
    z = np.zeros(1 << n)
    z[0] = 1
    a = run_gpu(n, ops, ud, z.copy(), np.zeros(1 << n))
    b = run_gpu(n, ops, ud, z.copy(), np.zeros(1 << n))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
