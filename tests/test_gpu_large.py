"""GPU: the largest state one MI355X holds (33 qubits: 128 GiB in complex128,
64 GiB in complex64) through the tile, block and bit-reversal kernels.

No CPU oracle runs at this size, so the check is the value-encoding known
answer: H on every qubit, the phases of a seeded 33-bit integer v and the
inverse QFT with bit reversal give exactly the basis state |rev_33(v)>.  A
wrong index anywhere in 2^33 amplitudes (a 32-bit overflow, a wrapped tile
or block number) moves probability away from that one index.  Measuring
some of its qubits then returns their bits with probability 1."""
import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd.state import GpuState, Plan

pytestmark = pytest.mark.gpu
N = 33


def big_state(dtype):
    try:
        return GpuState(N, dtype)
    except _lib.BflyError as e:   # a device with less memory than 2 * 2^33 amplitudes
        pytest.skip(f"{N}-qubit state does not fit: {e}")


@pytest.mark.parametrize("dtype,mode,tol", [
    (np.float64, _lib.MODE_AUTO, 1e-12),    # dense FP64 blocks (MFMA) for this circuit
    (np.float64, _lib.MODE_FUSED, 1e-12),   # tile passes
    (np.float32, _lib.MODE_AUTO, 1e-5),     # FP32 MFMA blocks
])
def test_33_qubit_value_encoding(gpu_lib, dtype, mode, tol):
    qc, v = W.value_encoding_precision(N, 2033)
    ops, ud = lower_circuit(qc)
    rev = int(format(v, f"0{N}b")[::-1], 2)
    print(f"[large] n={N} dtype={np.dtype(dtype).name} mode={mode} v={v}", flush=True)
    g = big_state(dtype)
    p = Plan(N, dtype, ops, ud, _lib.opts(mode=mode))
    p.execute(g)
    g.synchronize()
    pr = g.prob_mask((1 << N) - 1, rev)
    nrm = g.norm2()
    # measuring three qubits of the basis state returns their bits of rev with
    # probability 1 (outcome and collapse kernels over all 2^33 amplitudes)
    targets = [0, N - 1, 17]
    outcome, prob = g.measure(targets, None, 12345)
    nrm_after = g.norm2()
    p.free()
    g.free()
    assert pr > 1 - tol, 1 - pr
    assert abs(nrm - 1) < tol, nrm
    assert outcome == sum(((rev >> t) & 1) << k for k, t in enumerate(targets)), outcome
    assert abs(prob - 1) < tol and abs(nrm_after - 1) < tol, (prob, nrm_after)
