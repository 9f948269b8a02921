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
        if e.code != _lib.BFLY_ENOMEM:
            raise
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


def _hip():
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so.7")   # the runtime libbfly_hip.so is linked against
    hip.hipMemcpy.restype = ct.c_int
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    return hip


def _read(hip, base: int, idx, dtype) -> np.ndarray:
    """Amplitudes at `idx` of one device array (hipMemcpy, device to host)."""
    out = np.empty(len(idx), dtype=dtype)
    sz = np.dtype(dtype).itemsize
    for k, i in enumerate(idx):
        assert hip.hipMemcpy(out[k:].ctypes.data, base + int(i) * sz, sz, 2) == 0
    return out


@pytest.mark.parametrize("n,dtype", [(31, np.float64), (N, np.float32)])
def test_full_bit_reversal_beyond_one_block_per_tile(gpu_lib, n, dtype):
    """The tiled bit reversal loops over its middle indices once they exceed
    2^20 blocks (complex128 above 30 qubits, complex64 above 32): sampled
    amplitudes land at the reversed index, bit-exact."""
    import ctypes as ct
    try:
        g = GpuState(n, dtype)
    except _lib.BflyError as e:
        if e.code != _lib.BFLY_ENOMEM:
            raise
        pytest.skip(f"{n}-qubit state does not fit: {e}")
    print(f"[large] bitrev n={n} dtype={np.dtype(dtype).name}", flush=True)
    g.fill_random(4242)
    g.synchronize()
    pr, pi = ct.c_void_p(), ct.c_void_p()
    _lib.check(_lib.load().bfly_state_device_ptrs(g.handle, ct.byref(pr), ct.byref(pi)))
    hip = _hip()
    rng = np.random.default_rng(n)
    size = 1 << n
    idx = np.concatenate([[0, 1, 2, size - 1, size - 2, 1 << (n - 1), (1 << (n // 2)) - 1],
                          rng.integers(0, size, 2000, dtype=np.int64)]).astype(np.int64)
    rev = np.array([int(format(int(i), f"0{n}b")[::-1], 2) for i in idx], dtype=np.int64)
    before = (_read(hip, pr.value, idx, dtype), _read(hip, pi.value, idx, dtype))
    g.bitrev()
    g.synchronize()
    after = (_read(hip, pr.value, rev, dtype), _read(hip, pi.value, rev, dtype))
    g.free()
    for b, a in zip(before, after):
        assert np.array_equal(b.view(np.uint8), a.view(np.uint8))
