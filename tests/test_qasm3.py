"""OpenQASM 3 import (no GPU): the reference's round-trip test
(tests/test_qasm3_import_value_encoding.mojo: export encode_value_circuit(3,
4.7) as h / p / cp lines, re-import, states equal), the reference's angle
forms against their closed forms, the reference's error messages, and the
library's C++ front end (bfly_qasm3_lower) against the Python mirror: the same
op array, bit for bit, on seeded random programs."""
import math
import re

import numpy as np
import pytest

from code_mojo_amd import _lib, algos, lower_circuit, parse_qasm3_string
from code_mojo_amd.circuit import GateTransformation
from code_mojo_amd.gates import GateKind
from code_mojo_amd.lowering import LoweringError
from code_mojo_amd.qasm3 import QasmError, parse_angle
from oracle import oracle as O


def qasm_from_circuit(circuit) -> str:
    """The reference test's exporter: h, p(a), cp(a) with the angle's repr."""
    lines = ["OPENQASM 3;", 'include "stdgates.inc";', f"qubit[{circuit.num_qubits}] q;"]
    for tr in circuit.transformations:
        assert isinstance(tr, GateTransformation)
        k = tr.gate_info.kind
        if not tr.controls and k == GateKind.H:
            lines.append(f"h q[{tr.target}];")
        elif not tr.controls and k == GateKind.P:
            lines.append(f"p({tr.gate_info.arg!r}) q[{tr.target}];")
        elif len(tr.controls) == 1 and k == GateKind.P:
            lines.append(f"cp({tr.gate_info.arg!r}) q[{tr.controls[0]}], q[{tr.target}];")
        else:
            raise AssertionError("unsupported gate for export")
    return "\n".join(lines) + "\n"


def run(n, ops, ud):
    re = np.zeros(1 << n)
    re[0] = 1
    im = np.zeros(1 << n)
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


@pytest.mark.parametrize("n,v", [(3, 4.7), (6, 37.0), (8, 200.25)])
def test_reference_round_trip(n, v):
    qc = algos.encode_value_circuit(n, v)
    text = qasm_from_circuit(qc)
    imported = parse_qasm3_string(text)
    a = run(n, *lower_circuit(qc))
    b = run(n, *lower_circuit(imported))
    assert np.abs(a - b).max() <= 1e-10
    m, ops, ud = _lib.qasm3_lower(text)
    assert m == n
    assert np.abs(run(n, ops, ud) - a).max() <= 1e-10


@pytest.mark.parametrize("text,want", [
    ("pi", math.pi), ("-pi", -math.pi), ("3*pi", 3 * math.pi), ("-0.5 * pi", -0.5 * math.pi),
    ("pi/8", math.pi / 8), ("-pi/8", -math.pi / 8), ("2/pi", 2 / math.pi), ("0.125", 0.125),
    ("1e-3", 1e-3),
    # forms the reference rejects, evaluated left to right
    ("2*pi/3", 2 * math.pi / 3), ("-(pi/4) + 1", -(math.pi / 4) + 1), ("pi/2 - pi/4", math.pi / 2 - math.pi / 4),
])
def test_angle_forms(text, want):
    assert parse_angle(text) == want
    _, ops, _ = _lib.qasm3_lower(f"qubit[1] q;\np({text}) q[0];\n")
    assert ops[0]["theta"] == want


@pytest.mark.parametrize("text,msg", [
    ("h q[0];", "QASM missing qubit declaration before gates."),
    ("OPENQASM 3;\n", "No qubit declaration found in QASM."),
    ("qubit[2] q;\nfoo q[0];", "Unsupported QASM line: foo q[0]"),
    ("qubit[2] q;\ncp(pi) q[0];", "Invalid cp operands: q[0]"),
    ("qubit[2] q;\nh q;", "Invalid qubit token: q"),
    ("qubit[2 q;", "Invalid qubit declaration"),
    ("qubit[2] q;\np(2*pa) q[0];", "Invalid angle"),
])
def test_errors_match_in_both_front_ends(text, msg):
    with pytest.raises(QasmError, match=re.escape(msg)):
        parse_qasm3_string(text)
    with pytest.raises(_lib.BflyError, match=re.escape(msg)):
        _lib.qasm3_lower(text)


def test_out_of_range_qubits():
    text = "qubit[3] q;\nh q[5];\n"
    with pytest.raises(LoweringError, match="Target out of bounds: 5"):
        lower_circuit(parse_qasm3_string(text))
    with pytest.raises(_lib.BflyError, match="Target out of bounds: 5"):
        _lib.qasm3_lower(text)


GATES1 = ["h", "x", "y", "z", "s", "sdg", "t", "tdg"]
GATES1P = ["p", "phase", "u1", "rx", "ry", "rz"]
GATES2 = ["cx", "CX", "cnot", "cy", "cz"]
GATES2P = ["cp", "cphase", "cu1", "crx", "cry", "crz"]


def random_program(rng, regs):
    n = sum(regs.values())
    names = list(regs)

    def qref():
        r = names[int(rng.integers(len(names)))]
        return f"{r}[{int(rng.integers(regs[r]))}]"

    def distinct(k):
        while True:
            qs = [qref() for _ in range(k)]
            if len(set(qs)) == k:
                return qs

    def angle():
        forms = [lambda: repr(float(rng.uniform(-7, 7))), lambda: "pi", lambda: "-pi/%d" % rng.integers(1, 64),
                 lambda: "%d*pi" % rng.integers(-5, 6), lambda: "%g/pi" % rng.uniform(0.1, 3),
                 lambda: "2*pi/%d + 0.5" % rng.integers(1, 9)]
        return forms[int(rng.integers(len(forms)))]()

    lines = ["OPENQASM 3.0;", 'include "stdgates.inc";']
    lines += [f"qubit[{s}] {r};" for r, s in regs.items()]
    lines.append(f"bit[{n}] c;")
    for _ in range(60):
        kind = int(rng.integers(9))
        if kind == 0:
            lines.append(f"{GATES1[int(rng.integers(len(GATES1)))]} {qref()};")
        elif kind == 1:
            lines.append(f"{GATES1P[int(rng.integers(len(GATES1P)))]}({angle()}) {qref()};")
        elif kind == 2:
            a, b = distinct(2)
            lines.append(f"{GATES2[int(rng.integers(len(GATES2)))]} {a}, {b};")
        elif kind == 3:
            a, b = distinct(2)
            lines.append(f"{GATES2P[int(rng.integers(len(GATES2P)))]}({angle()}) {a}, {b};")
        elif kind == 4:
            lines.append(f"u3({angle()}, {angle()}, {angle()}) {qref()}; u2({angle()}, {angle()}) {qref()};")
        elif kind == 5 and n >= 3:
            a, b, c = distinct(3)
            lines.append(f"ccx {a}, {b}, {c};  // Toffoli")
        elif kind == 6 and n >= 3:
            k = int(rng.integers(2, min(n, 5) + 1))
            lines.append(f"mcp({angle()}) {', '.join(distinct(k))};")
            lines.append(f"ctrl({k - 1}) @ ry({angle()}) {', '.join(distinct(k))};")
        elif kind == 7:
            a, b = distinct(2)
            lines.append(f"swap {a}, {b}; barrier {a};")
        else:
            lines.append(f"c[0] = measure {qref()};")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed", range(12))
def test_cpp_front_end_equals_python_mirror(seed):
    rng = np.random.default_rng(seed)
    regs = {"q": int(rng.integers(3, 9))} if seed % 2 == 0 else {"a": 3, "b": int(rng.integers(2, 6)), "anc": 1}
    text = random_program(rng, regs)
    n, ops, ud = _lib.qasm3_lower(text)
    qc = parse_qasm3_string(text)
    ops_py, ud_py = lower_circuit(qc)
    assert n == qc.num_qubits == sum(regs.values())
    assert ops.tobytes() == ops_py.tobytes()
    assert ud.tobytes() == ud_py.tobytes()


def test_registers_broadcast_measure_and_statements_per_line():
    text = "qubit[2] a; qubit[3] b; h a[1]; cx a[1], b[2]; bit[3] c = measure b;\n"
    n, ops, ud = _lib.qasm3_lower(text)
    assert n == 5
    assert ops[0]["target"] == 1 and ops[1]["target"] == 4 and ops[1]["controls"] == 1 << 1
    assert ops[2]["kind"] == 6 and ops[2]["arg_i"] == 3 and list(ud[0::2]) == [2, 3, 4]
