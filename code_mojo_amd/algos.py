"""Circuit builders of the reference's algorithms (butterfly/algos/*).

Each builder produces the same transformation list as the reference's Mojo
builder, so a circuit built here and run with the GPU strategy is the
circuit the reference would run with its CPU strategies.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

from .circuit import (Circuit, ClassicalReplacementKind, ClassicalTransformation, QuantumCircuit,
                      QuantumRegister, Register, bit_reverse, replace_quantum_with_classical)
from .gates import pi


# -- value encoding / QFT (algos/value_encoding_circuit.mojo:5-29) ------------
def iqft_circuit(qc: Circuit, targets: Sequence[int], swap: bool = False) -> None:
    """Inverse QFT on `targets` (value_encoding_circuit.mojo:5-12)."""
    for j in reversed(range(len(targets))):
        qc.h(targets[j])
        for k in reversed(range(j)):
            qc.cp(targets[j], targets[k], -pi / 2 ** (j - k))
    if swap:
        bit_reverse(qc, list(targets))


def build_iqft_circuit(n: int, swap: bool = False) -> Circuit:
    qc = QuantumCircuit(n)
    iqft_circuit(qc, [n - 1 - j for j in range(n)], swap=swap)
    return qc


def qft_circuit(n: int, targets: Optional[Sequence[int]] = None, swap: bool = True) -> Circuit:
    """Forward QFT as the inverse of iqft_circuit (tests/test_qft_classical.mojo:40-56)."""
    qc = QuantumCircuit(n)
    iqft_circuit(qc, list(targets) if targets is not None else [n - 1 - j for j in range(n)], swap=swap)
    return qc.inverse()


def value_encoding_angle(j: int, v, exact: bool = False) -> float:
    """Angle of qubit j: 2*pi/2^(j+1)*v (value_encoding_circuit.mojo:27).

    With exact=True and integer v, v is first reduced modulo 2^(j+1) in
    integers, so the FP64 angle carries no error from the size of v."""
    if exact:
        if int(v) != v:
            raise ValueError("exact angle generation needs an integer value")
        r = int(v) % (1 << (j + 1))
        return 2 * pi * r / 2 ** (j + 1)
    return 2 * pi / 2 ** (j + 1) * v


def encode_value_circuit(n: int, v, exact_angles: bool = False, swap: bool = False) -> Circuit:
    """H on all qubits, P(j, 2pi/2^(j+1) v), IQFT on [n-1..0]
    (value_encoding_circuit.mojo:19-29).  For integer v the output is the
    basis state |v> (|rev_n(v)> with swap=True)."""
    qc = QuantumCircuit(n)
    for j in range(n):
        qc.h(j)
    for j in range(n):
        qc.p(j, value_encoding_angle(j, v, exact_angles))
    iqft_circuit(qc, [n - 1 - j for j in range(n)], swap=swap)
    return qc


# -- Grover (algos/grover.mojo:21-195) ----------------------------------------
def inversion_0_circuit(n: int) -> Circuit:
    """X(all), MCP(pi), X(all) (grover.mojo:21-42)."""
    q = QuantumRegister("q", n)
    qc = QuantumCircuit(q)
    for i in range(n):
        qc.x(i)
    qc.mcp(list(range(n - 1)), n - 1, pi)
    for i in range(n):
        qc.x(i)
    return qc


def diffuser_circuit(n: int) -> Circuit:
    """H(all), inversion_0, H(all) (grover.mojo:45-62)."""
    q = QuantumRegister("q", n)
    qc = QuantumCircuit(q)
    for i in range(n):
        qc.h(i)
    qc.append_circuit(inversion_0_circuit(n), q)
    for i in range(n):
        qc.h(i)
    return qc


def phase_oracle_match(n: int, items: Sequence[int]) -> Circuit:
    """X on zero bits, MCP(pi), X on zero bits, per item (grover.mojo:68-94)."""
    q = QuantumRegister("q", n)
    qc = QuantumCircuit(q)
    for m in items:
        for i in range(n):
            if (m >> i) & 1 == 0:
                qc.x(i)
        qc.mcp(list(range(n - 1)), n - 1, pi)
        for i in range(n):
            if (m >> i) & 1 == 0:
                qc.x(i)
    return qc


def grover_iterate_circuit(oracle, num_qubits: int) -> Circuit:
    """oracle + diffuser (grover.mojo:107-119)."""
    qc = QuantumCircuit(num_qubits)
    if isinstance(oracle, Circuit):
        qc.append_circuit(oracle, QuantumRegister("q", num_qubits))
    else:
        qc.add_classical(oracle.name, oracle.targets, oracle.apply)
    qc.append_circuit(diffuser_circuit(num_qubits), QuantumRegister("q", num_qubits))
    return qc


def grover_iterate_circuit_classical_diffuser(oracle, num_qubits: int) -> Circuit:
    """grover_iterate_circuit with the diffuser rewritten into the classical
    GROVER_DIFFUSER transform (grover.mojo:122-133)."""
    qc = grover_iterate_circuit(oracle, num_qubits)
    replace_quantum_with_classical(qc, [ClassicalReplacementKind.GROVER_DIFFUSER])
    return qc


def append_grover_to_register(qc: Circuit, register: Register, oracle, iterations: int) -> None:
    """Oracle + diffuser `iterations` times on `register` (grover.mojo:167-180)."""
    step = grover_iterate_circuit(oracle, register.length)
    for _ in range(iterations):
        qc.append_circuit(step, register)


def append_grover_to_register_classical_diffuser(qc: Circuit, register: Register, oracle,
                                                 iterations: int) -> None:
    """The same with the classical diffuser (grover.mojo:183-197)."""
    step = grover_iterate_circuit_classical_diffuser(oracle, register.length)
    for _ in range(iterations):
        qc.append_circuit(step, register)


def grover_circuit(items: Sequence[int], num_qubits: int, iterations: int, use_shortcut: bool = False) -> Circuit:
    """Grover iterations without the initial H layer (grover.mojo:135-164)."""
    if use_shortcut:
        oracle = ClassicalTransformation("sign_flip_oracle", list(items))
    else:
        oracle = phase_oracle_match(num_qubits, items)
    q = QuantumRegister("q", num_qubits)
    qc = QuantumCircuit(q)
    step = grover_iterate_circuit(oracle, num_qubits)
    for _ in range(iterations):
        qc.append_circuit(step, q)
    return qc


def grover_search_circuit(n: int, marked: int, iterations: Optional[int] = None) -> Circuit:
    """H layer + floor(pi/4 sqrt(2^n)) Grover iterations on one marked state."""
    if iterations is None:
        iterations = int(math.floor(pi / 4 * math.sqrt(2 ** n)))
    q = QuantumRegister("q", n)
    qc = QuantumCircuit(q)
    for i in range(n):
        qc.h(i)
    qc.append_circuit(grover_circuit([marked], n, iterations), q)
    return qc


# -- amplitude estimation (algos/amplitude_estimation.mojo:5-76) --------------
def append_amplitude_estimation(qc: Circuit, c: Register, q: Register, prepare: Circuit, oracle: Circuit,
                                swap: bool = True) -> None:
    n = c.length
    qc.append_circuit(prepare, q)
    for i in range(n):
        qc.h(c[i])
    n_state = prepare.num_qubits
    grover_op = QuantumCircuit(n_state)
    q_dummy = QuantumRegister("q", n_state)
    grover_op.append_circuit(oracle, q_dummy)
    grover_op.append_circuit(prepare.inverse(), q_dummy)
    grover_op.append_circuit(inversion_0_circuit(n_state), q_dummy)
    grover_op.append_circuit(prepare, q_dummy)
    for i in range(n):
        ctrl = c[i] if swap else c[n - 1 - i]
        for _ in range(1 << i):
            qc.c_append_circuit(ctrl, grover_op, q)
    iqft = QuantumCircuit(n)
    iqft_circuit(iqft, [n - 1 - j for j in range(n)], swap=False)
    qc.append_circuit(iqft, c)


def amplitude_estimation_circuit(n: int, prepare: Circuit, oracle: Circuit, swap: bool = True) -> Circuit:
    total = n + prepare.num_qubits
    qc = QuantumCircuit(total)
    c = QuantumRegister("c", n)
    # The reference builds the state register with Register's default
    # start = 0 (amplitude_estimation.mojo:70-72), so it overlaps c[0]; gates
    # whose control lands on their own target never fire.  Reproduced as is.
    q = QuantumRegister("q", prepare.num_qubits)
    append_amplitude_estimation(qc, c, q, prepare, oracle, swap)
    return qc


def ae_estimate_from_peak(y: int, m: int) -> float:
    """a = cos^2(pi y / 2^m): the reference's Grover operator carries a global
    phase of pi (tests/test_amplitude_estimation.mojo:96-102)."""
    return math.cos(pi * y / (1 << m)) ** 2
