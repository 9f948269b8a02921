"""Dense unitary blocks of 4 to 10 qubits (UnitaryTransformation /
ControlledUnitaryTransformation, transformations_unitary.mojo:22-139).

CPU: the oracle's UNITARY equals the dense 2^n x 2^n operator built with
Kronecker embedding (block index bit i = qubit target + i, the control
qubits selecting the columns it acts on) for m = 4 and 5, with and without
controls.

GPU: unitary_kernel (one 2^m group of amplitudes per workgroup slice, LDS
staged, m = 4..10) against the oracle on random states: complex128 within
1e-12 max |amplitude error|, complex64 within 2e-5 of the largest amplitude,
at the lowest, a middle and the highest target offset, with 0, 1 and 2
controls."""
import numpy as np
import pytest

from code_mojo_amd import _lib
from code_mojo_amd import workloads as W
from code_mojo_amd.lowering import OpList
from oracle import oracle as O


def haar(m, seed):
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(1 << m, 1 << m)) + 1j * rng.normal(size=(1 << m, 1 << m))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def dense(n, u, target, m, controls):
    """The 2^n operator of u on qubits [target, target + m) under `controls`."""
    N = 1 << n
    tmask = ((1 << m) - 1) << target
    cmask = sum(1 << c for c in controls)
    M = np.zeros((N, N), dtype=complex)
    for col in range(N):
        if (col & cmask) != cmask:
            M[col, col] = 1.0
            continue
        k = (col >> target) & ((1 << m) - 1)
        base = col & ~tmask
        for r in range(1 << m):
            M[base | (r << target), col] += u[r, k]
    return M


def oracle_run(n, ops, ud, psi):
    re, im = psi.real.copy(), psi.imag.copy()
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


@pytest.mark.parametrize("m,target,controls", [(4, 0, []), (4, 2, [1]), (5, 1, [0, 6]), (5, 2, [])])
def test_oracle_unitary_against_dense_embedding(m, target, controls):
    n = 7
    u = haar(m, 10 * m + target)
    out = OpList(n)
    out.unitary(u, target, m, controls)
    ops, ud = out.finish()
    rng = np.random.default_rng(m)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    want = dense(n, u, target, m, controls) @ psi
    assert np.abs(oracle_run(n, ops, ud, psi) - want).max() < 1e-13


CASES = []
for m in range(4, 11):
    n = 14
    for target, controls in ((0, []), ((n - m) // 2, [n - 1] if (n - m) // 2 + m <= n - 1 else []),
                             (n - m, [0, 1] if n - m >= 2 else [])):
        CASES.append((n, m, target, controls))


@pytest.mark.gpu
@pytest.mark.parametrize("n,m,target,controls", CASES)
def test_unitary_kernel_widths_vs_oracle(gpu_lib, n, m, target, controls):
    from code_mojo_amd.state import GpuState, Plan
    u = haar(m, 100 * m + target)
    out = OpList(n)
    out.unitary(u, target, m, controls)
    out.unitary(haar(m, 7 + m).conj().T @ u, target, m, controls)     # two in a row
    ops, ud = out.finish()
    st = W.normal_state(n, m + target)
    ref = oracle_run(n, ops, ud, st.re + 1j * st.im)
    scale = float(np.abs(ref).max())
    for dtype, tol in ((np.float64, 1e-12), (np.float32, 2e-5 * scale)):
        g = GpuState(n, dtype)
        g.upload_arrays(st.re.astype(dtype), st.im.astype(dtype))
        p = Plan(n, dtype, ops, ud, _lib.opts())
        p.execute(g)
        got = g.download()
        p.free()
        g.free()
        err = float(np.abs(got.re.astype(np.float64) + 1j * got.im.astype(np.float64) - ref).max())
        assert err <= tol, (m, target, controls, np.dtype(dtype).name, err)
