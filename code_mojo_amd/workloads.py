"""Seeded synthetic workloads of the benchmark configurations.

  qft_random_state     n-qubit QFT (H + controlled-phase ladder pi/2^k + full
                       bit reversal) on a state with re/im ~ N(0,1), normalised
  random_layered       layers of U3 on every qubit, CZ on a brick of adjacent
                       pairs (offset alternating 0/1), one CX on a random pair
  value_encoding       H on all qubits, value-encoding phases, inverse QFT with
                       bit reversal (exact dyadic angles for integer values)
  grover_search        H layer + floor(pi/4 sqrt(2^n)) iterations, one mark
  amplitude_estimation 8-qubit phase register, Ry(2 asin sqrt a) preparation

Random numbers: a counter-based SplitMix64 stream (the same integer sequence
in any language) turned into uniforms with 53-bit mantissas; normal deviates
by Box-Muller in FP64.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from . import algos
from .circuit import Circuit, QuantumCircuit, bit_reverse
from .gates import Custom_Gate, U3, pi
from .state import QuantumState

_M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """`count` outputs of SplitMix64 started at `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniforms(seed: int, count: int) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal_state(n: int, seed: int, dtype=np.float64) -> QuantumState:
    """re, im i.i.d. N(0,1) (Box-Muller on SplitMix64 uniforms), normalised
    with a compensated (math.fsum) norm."""
    size = 1 << n
    u = uniforms(seed, 2 * size)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    re = r * np.cos(2 * pi * u2)
    im = r * np.sin(2 * pi * u2)
    norm = math.sqrt(math.fsum(np.concatenate([re * re, im * im]).tolist()) if size <= (1 << 16)
                     else float(np.sum(re * re, dtype=np.float64) + np.sum(im * im, dtype=np.float64)))
    return QuantumState(n, dtype, (re / norm).astype(dtype), (im / norm).astype(dtype))


def qft_circuit(n: int) -> Circuit:
    """For j = n-1..0: H(j), then CP(k -> j, pi/2^(j-k)) for k < j; then the
    full bit reversal.  Equivalent to the forward QFT built as the inverse of
    iqft_circuit (algos/value_encoding_circuit.mojo:5-12)."""
    qc = QuantumCircuit(n)
    for j in range(n - 1, -1, -1):
        qc.h(j)
        for k in range(j - 1, -1, -1):
            qc.cp(k, j, pi / 2 ** (j - k))
    bit_reverse(qc)
    return qc


def random_layered(n: int, layers: int, seed: int) -> Circuit:
    """Layered U3 + CZ-brick + one CX circuit (angles ~ U[0, 2pi))."""
    qc = QuantumCircuit(n)
    u = uniforms(seed, layers * (3 * n + 2))
    pos = 0
    for layer in range(layers):
        for q in range(n):
            th, ph, la = (2 * pi * u[pos], 2 * pi * u[pos + 1], 2 * pi * u[pos + 2])
            pos += 3
            qc.add(q, Custom_Gate(U3(th, ph, la), "U3"))
        start = layer % 2
        for a in range(start, n - 1, 2):
            qc.cz(a, a + 1)
        c = int(u[pos] * n) % n
        t = int(u[pos + 1] * (n - 1)) % (n - 1)
        if t >= c:
            t += 1
        pos += 2
        qc.cx(c, t)
    return qc


def value_encoding_precision(n: int, seed: int) -> Tuple[Circuit, int]:
    """H on all qubits, value-encoding phases for a seeded n-bit integer v
    (exact dyadic FP64 angles), inverse QFT with bit reversal.  The exact
    output is the basis state |rev_n(v)>."""
    v = int(splitmix64(seed, 1)[0]) & ((1 << n) - 1)
    return algos.encode_value_circuit(n, v, exact_angles=True, swap=True), v


def grover_search(n: int, seed: int) -> Tuple[Circuit, int]:
    marked = int(splitmix64(seed, 1)[0] % np.uint64(1 << n))
    return algos.grover_search_circuit(n, marked), marked


def amplitude_estimation(seed: int, eval_qubits: int = 8) -> Tuple[Circuit, float]:
    a = float(uniforms(seed, 1)[0])
    prep = QuantumCircuit(1)
    prep.ry(0, 2 * math.asin(math.sqrt(a)))
    oracle = QuantumCircuit(1)
    oracle.z(0)
    return algos.amplitude_estimation_circuit(eval_qubits, prep, oracle), a


def layered_gate_layers(layers: int) -> int:
    return layers


def count_gates(circuit: Circuit) -> int:
    return len(circuit.transformations)


def loguniform_qubits(seed: int, count: int, lo: int, hi: int) -> List[int]:
    """Qubit counts in [lo, hi], log-uniform (uniform in log n)."""
    u = uniforms(seed, count)
    vals = np.exp(np.log(lo) + u * (np.log(hi + 1) - np.log(lo)))
    return [min(hi, max(lo, int(v))) for v in vals]
