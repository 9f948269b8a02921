"""The oracle (CPU restatement of the reference's scalar strategy) pinned
against closed forms and the reference's own known-answer assertions.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign, a transposed operand, a wrong bit) fails it; the last
test plants such mistakes in a copy of the op list and confirms they are
caught."""
import math

import numpy as np
import pytest

from code_mojo_amd import algos, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd.circuit import QuantumCircuit, QuantumRegister, bit_reverse
from code_mojo_amd.lowering import OP_DTYPE
from oracle import oracle as O

from .helpers import GOLDEN


def run(n, circuit_or_ops, re=None, im=None, dtype=np.float64):
    if isinstance(circuit_or_ops, tuple):
        ops, ud = circuit_or_ops
    else:
        ops, ud = lower_circuit(circuit_or_ops)
    if re is None:
        re = np.zeros(1 << n, dtype=dtype)
        im = np.zeros(1 << n, dtype=dtype)
        re[0] = 1
    re = np.array(re, dtype=dtype, copy=True)
    im = np.array(im, dtype=dtype, copy=True)
    O.execute(n, re, im, ops, ud)
    return re, im


def test_dense_matrix_fixtures():
    """Every gate kind, 0..3 controls, dense blocks, swaps, bit reversals vs
    the Kronecker-product matrix of the same op list."""
    d = np.load(GOLDEN / "dense_small.npz")
    assert int(d["count"]) >= 30
    for i in range(int(d["count"])):
        n = int(d[f"c{i}_n"])
        ops = d[f"c{i}_ops"].view(OP_DTYPE)
        psi = d[f"c{i}_in"]
        re, im = run(n, (ops, d[f"c{i}_udata"]), psi.real, psi.imag)
        assert np.abs(re + 1j * im - d[f"c{i}_out"]).max() < 1e-14, i


def test_qft_is_scaled_inverse_fft():
    d = np.load(GOLDEN / "qft_fft.npz")
    for n in range(2, 13):
        psi = d[f"n{n}_in"]
        re, im = run(n, W.qft_circuit(n), psi.real, psi.imag)
        assert np.abs(re + 1j * im - d[f"n{n}_out"]).max() < 1e-13, n


@pytest.mark.parametrize("n,v", [(3, 5), (5, 11), (6, 37), (8, 200), (10, 777)])
def test_value_encoding_integer_is_basis_state(n, v):
    re, im = run(n, algos.encode_value_circuit(n, v))
    p = re ** 2 + im ** 2
    assert int(np.argmax(p)) == v and abs(p[v] - 1) < 1e-13


def test_value_encoding_fejer_peak():
    """n=3, v=4.7: peak 0.740258 at index 5 (Fejer kernel)."""
    re, im = run(3, algos.encode_value_circuit(3, 4.7))
    p = re ** 2 + im ** 2
    assert int(np.argmax(p)) == 5 and abs(p[5] - 0.740258) < 5e-7


def test_value_encoding_exact_angles_with_swap():
    n, v = 12, 3001
    re, im = run(n, algos.encode_value_circuit(n, v, exact_angles=True, swap=True))
    rev = int(format(v, f"0{n}b")[::-1], 2)
    assert abs(re[rev] ** 2 + im[rev] ** 2 - 1) < 1e-13


def test_bitrev_tables_and_involution():
    """Full and partial reversal == integer tables; involution; full == partial
    over all targets (tests/test_bit_reverse.mojo:23-58)."""
    d = np.load(GOLDEN / "bitrev.npz")
    for n in range(1, 13):
        st = W.normal_state(n, n)
        re, im = st.re.copy(), st.im.copy()
        O.bitrev(n, re, im, 0)
        assert np.array_equal(re, st.re[d[f"full_{n}"]])
        re2, im2 = st.re.copy(), st.im.copy()
        O.bitrev(n, re2, im2, (1 << n) - 1)
        assert np.array_equal(re, re2) and np.array_equal(im, im2)
        O.bitrev(n, re, im, 0)
        assert np.array_equal(re, st.re) and np.array_equal(im, st.im)
        if n >= 2:
            mask = int(d[f"mask_{n}"])
            re, im = st.re.copy(), st.im.copy()
            O.bitrev(n, re, im, mask)
            assert np.array_equal(re, st.re[d[f"partial_{n}"]])
            O.bitrev(n, re, im, mask)
            assert np.array_equal(re, st.re)


def test_grover_closed_form():
    """P(marked) after k iterations = sin^2((2k+1) theta)
    (algos/grover_classical.mojo:154-168); reference threshold: n=3, 2 its,
    P >= 0.9 (tests/test_grover.mojo:18-44)."""
    d = np.load(GOLDEN / "grover.npz")
    for n in range(3, 9):
        marked = (5 * n + 3) % (1 << n)
        for k in range(len(d[f"n{n}_p"])):
            re, im = run(n, algos.grover_search_circuit(n, marked, k))
            assert abs(re[marked] ** 2 + im[marked] ** 2 - d[f"n{n}_p"][k]) < 1e-12, (n, k)
    q = QuantumRegister("q", 3)
    qc = QuantumCircuit(q)
    for i in range(3):
        qc.h(i)
    qc.append_circuit(algos.grover_circuit([7], 3, 2), q)
    re, im = run(3, qc)
    assert re[7] ** 2 + im[7] ** 2 >= 0.9


def test_grover_shortcut_equals_gate_oracle():
    n, marked = 5, 19
    a = run(n, algos.grover_search_circuit(n, marked, 3))
    q = QuantumRegister("q", n)
    qc = QuantumCircuit(q)
    for i in range(n):
        qc.h(i)
    qc.append_circuit(algos.grover_circuit([marked], n, 3, use_shortcut=True), q)
    b = run(n, qc)
    assert np.abs(a[0] - b[0]).max() < 1e-14 and np.abs(a[1] - b[1]).max() < 1e-14


def test_amplitude_estimation_reference_case():
    """theta = pi/3 (a = 0.25), 5 evaluation qubits: |a_hat - a| <= 0.1 with
    a_hat = cos^2(pi y / 2^m) (tests/test_amplitude_estimation.mojo:78-180)."""
    prep = QuantumCircuit(1)
    prep.ry(0, math.pi / 3)
    orc = QuantumCircuit(1)
    orc.z(0)
    qc = algos.amplitude_estimation_circuit(5, prep, orc)
    n = qc.num_qubits
    re, im = run(n, qc)
    best, best_p = 0, -1.0
    for y in range(32):
        p = re[y] ** 2 + im[y] ** 2 + re[y + 32] ** 2 + im[y + 32] ** 2
        if p > best_p:
            best, best_p = y, p
    assert abs(algos.ae_estimate_from_peak(best, 5) - 0.25) <= 0.1


def test_qft_gate_vs_classical_orderings():
    """Gate IQFT with targets [3,2,1,0] and swap=False leaves the output
    bit-reversed relative to the FFT, and swap=True undoes it
    (algos/qft_classical.mojo:137-153 semantics)."""
    n = 4
    st = W.normal_state(n, 3)
    psi = st.to_complex()
    for swap in (False, True):
        qc = QuantumCircuit(n)
        algos.iqft_circuit(qc, [3, 2, 1, 0], swap=swap)
        re, im = run(n, qc, st.re, st.im)
        # targets MSB->LSB: input bit-reversed, FFT (forward sign for IQFT is -)
        rev = np.array([int(format(i, "04b")[::-1], 2) for i in range(16)])
        ref = np.fft.fft(psi[rev]) / 4.0
        if swap:
            ref = ref[rev]
        assert np.abs(re + 1j * im - ref).max() < 1e-13, swap


def test_unitary_round_trip():
    """U U^-1 |0> = |0> (tests/test_circuit_inverse.mojo:104-167 pattern)."""
    n = 6
    qc = QuantumCircuit(n)
    for j in range(n):
        qc.h(j)
        qc.rx(j, 0.3 + j)
        qc.ry((j + 1) % n, 1.1 * j)
    qc.cp(0, 5, 0.7)
    qc.crz(2, 4, -1.3)
    qc.ccx(0, 1, 3)
    bit_reverse(qc, [1, 3, 4])
    full = QuantumCircuit(n)
    full.append_circuit(qc)
    full.append_circuit(qc.inverse())
    re, im = run(n, full)
    assert abs(re[0] - 1) < 1e-13 and np.abs(re[1:]).max() < 1e-13 and np.abs(im).max() < 1e-13


def test_float32_oracle_close_to_float64():
    n = 10
    st = W.normal_state(n, 9)
    ops = lower_circuit(W.qft_circuit(n))
    a = run(n, ops, st.re, st.im)
    b = run(n, ops, st.re.astype(np.float32), st.im.astype(np.float32), dtype=np.float32)
    assert np.abs(a[0] - b[0]).max() < 1e-5


def test_planted_mistakes_are_caught():
    """Mutations of the op list (a sign, a transposed matrix, a wrong bit)
    must change the result beyond the parity bar -- the fixture checks are
    not vacuous."""
    d = np.load(GOLDEN / "dense_small.npz")
    i = next(k for k in range(int(d["count"])) if int(d[f"c{k}_n"]) == 5)
    n = 5
    ops = d[f"c{i}_ops"].view(OP_DTYPE).copy()
    psi = d[f"c{i}_in"]
    expected = d[f"c{i}_out"]
    gate_rows = np.nonzero((ops["kind"] == 0) & (ops["gate"] == 8))[0]
    assert gate_rows.size
    r = gate_rows[0]
    mutants = []
    m1 = ops.copy()
    m1["mat"][r, 3] = -m1["mat"][r, 3]          # sign of Im g01
    mutants.append(m1)
    m2 = ops.copy()
    m2["mat"][r, 2:4], m2["mat"][r, 4:6] = ops["mat"][r, 4:6].copy(), ops["mat"][r, 2:4].copy()  # transpose
    mutants.append(m2)
    m3 = ops.copy()
    m3["target"][r] = (m3["target"][r] + 1) % n  # wrong qubit
    if (int(m3["controls"][r]) >> int(m3["target"][r])) & 1:
        m3["target"][r] = (m3["target"][r] + 1) % n
    mutants.append(m3)
    for mu in mutants:
        try:
            re, im = run(n, (mu, d[f"c{i}_udata"]), psi.real, psi.imag)
        except Exception:
            continue
        assert np.abs(re + 1j * im - expected).max() > 1e-6
