"""OpenQASM 3 import (butterfly/utils/qasm3_import.mojo:7-123).

`parse_qasm3_string(text)` returns a QuantumCircuit.  The reference's
statements and angle forms (qubit[n], h, x, p(a), cp(a); angles pi, c*pi,
pi/d, c/pi and numbers) give its results and its error messages.  Also
accepted, exactly as the library's C++ front end `bfly_qasm3_lower`
(csrc/qasm.cpp) accepts them, so that both produce the same op array:
named registers, y z s sdg t tdg rx ry rz u3 U u2 u1 phase cx cy cz crx cry
crz ccx cphase cu1 swap mcp, the ctrl(k) @ modifier, measurements, barrier
and bit declarations (ignored), several statements per line, and arithmetic
angle expressions (+ - * / parentheses, pi).

Two deliberate differences from the reference: a second `qubit` declaration
adds a register (the reference starts a new circuit), and statements are
split on ';' (the reference reads one statement per line and ignores the
rest of a line).
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

from .circuit import MeasurementTransformation, QuantumCircuit
from .gates import Custom_Gate, GateInfo, GateKind, H_Gate, P_Gate, RX_Gate, RY_Gate, RZ_Gate, U3, X_Gate, Y_Gate, Z_Gate

pi = math.pi


class QasmError(ValueError):
    pass


def _number(t: str) -> float:
    s = t.strip()
    try:
        return float(s)
    except ValueError:
        raise QasmError(f"Invalid number: {s}") from None


class _Expr:
    """+ - * / unary minus, parentheses, pi, numbers (left-associative)."""

    def __init__(self, s: str):
        self.s, self.i = s, 0

    def ws(self):
        while self.i < len(self.s) and self.s[self.i].isspace():
            self.i += 1

    def primary(self) -> float:
        self.ws()
        s = self.s
        if self.i < len(s) and s[self.i] == "(":
            self.i += 1
            v = self.sum()
            self.ws()
            if self.i >= len(s) or s[self.i] != ")":
                raise QasmError(f"Invalid angle: {s}")
            self.i += 1
            return v
        if self.i < len(s) and s[self.i] == "-":
            self.i += 1
            return -self.primary()
        if self.i < len(s) and s[self.i] == "+":
            self.i += 1
            return self.primary()
        if s.startswith("pi", self.i):
            self.i += 2
            return pi
        b = self.i
        while self.i < len(s) and (s[self.i].isdigit() or s[self.i] in ".eE" or
                                   (s[self.i] in "+-" and self.i > b and s[self.i - 1] in "eE")):
            self.i += 1
        if self.i == b:
            raise QasmError(f"Invalid angle: {s}")
        return _number(s[b:self.i])

    def product(self) -> float:
        v = self.primary()
        while True:
            self.ws()
            if self.i < len(self.s) and self.s[self.i] == "*":
                self.i += 1
                v = v * self.primary()
            elif self.i < len(self.s) and self.s[self.i] == "/":
                self.i += 1
                v = v / self.primary()
            else:
                return v

    def sum(self) -> float:
        v = self.product()
        while True:
            self.ws()
            if self.i < len(self.s) and self.s[self.i] == "+":
                self.i += 1
                v = v + self.product()
            elif self.i < len(self.s) and self.s[self.i] == "-":
                self.i += 1
                v = v - self.product()
            else:
                return v


def parse_angle(expr: str) -> float:
    """parse_angle (qasm3_import.mojo:7-30), then the expression form."""
    trimmed = expr.strip()
    if "pi" in trimmed:
        s = trimmed.replace(" ", "")
        sign = 1.0
        if s.startswith("-"):
            sign = -1.0
            s = s[1:]
        try:
            if s == "pi":
                return sign * pi
            if s.endswith("*pi"):
                return sign * _number(s[:-3]) * pi
            if s.startswith("pi/"):
                return sign * pi / _number(s[3:])
            p = s.find("/pi")
            if p >= 0 and s.find("/pi", p + 1) < 0 and p + 3 == len(s):
                return sign * _number(s[:p]) / pi
        except QasmError:
            pass
    else:
        try:
            return _number(trimmed)
        except QasmError:
            pass
    e = _Expr(trimmed)
    v = e.sum()
    e.ws()
    if e.i != len(trimmed):
        raise QasmError(f"Invalid angle: {trimmed}")
    return v


def _qubit_index(regs: Dict[str, Tuple[int, int]], token: str) -> int:
    """parse_qubit_index (qasm3_import.mojo:43-50) plus the register offset."""
    t = token.strip()
    if t.endswith(";"):
        t = t[:-1].strip()
    if t.endswith(","):
        t = t[:-1].strip()
    lb, rb = t.find("["), t.find("]")
    if lb < 0 or rb < 0 or rb <= lb:
        raise QasmError(f"Invalid qubit token: {t}")
    name, idx = t[:lb].strip(), t[lb + 1:rb].strip()
    try:
        v = int(idx, 10)
    except ValueError:
        raise QasmError(f"Invalid qubit token: {t}") from None
    offset = 0
    if len(regs) > 1:
        if name not in regs:
            raise QasmError(f"Unknown register: {name}")
        offset = regs[name][0]
    return offset + v


def _split_args(inside: str) -> List[float]:
    if not inside.strip():
        return []
    out, cur, depth = [], "", 0
    for ch in inside:
        if ch == "(":
            depth += 1
        if ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(parse_angle(cur))
            cur = ""
        else:
            cur += ch
    out.append(parse_angle(cur))
    return out


def _split_operands(rest: str) -> List[str]:
    parts = rest.split(",")
    if len(parts) == 1 and not parts[0].strip():
        return []
    return [p.strip() for p in parts]


_FIXED = {"h": H_Gate, "x": X_Gate, "y": Y_Gate, "z": Z_Gate}
_ROT = {"rx": RX_Gate, "ry": RY_Gate, "rz": RZ_Gate}


def parse_qasm3_string(qasm: str) -> QuantumCircuit:
    regs: Dict[str, Tuple[int, int]] = {}
    n = 0
    has_circuit = False
    items: List[tuple] = []   # ("gate", controls, target, info) | ("swap", a, b) | ("measure", targets)

    def statement(s: str):
        nonlocal n, has_circuit
        if not s or s.startswith("OPENQASM") or s.startswith("include"):
            return
        if s.startswith("qubit") or s.startswith("qreg"):
            lb, rb = s.find("["), s.find("]")
            if lb < 0 or rb < 0:
                raise QasmError(f"Invalid qubit declaration: {s}")
            size = int(_number(s[lb + 1:rb]))
            name = s[4:lb].strip() if s.startswith("qreg") else s[rb + 1:].strip()
            if name in regs:
                raise QasmError(f"Register declared twice: {name}")
            regs[name] = (n, size)
            n += size
            has_circuit = True
            if n > 64:
                raise QasmError("more than 64 qubits")
            return
        m = s
        eq = m.find("=")
        if eq >= 0 and m[eq + 1:].strip().startswith("measure"):
            m = m[eq + 1:].strip()
        if m.startswith("measure ") or m.startswith("measure\t"):
            if not has_circuit:
                raise QasmError("QASM missing qubit declaration before gates.")
            rest = m[7:].strip()
            arrow = rest.find("->")
            if arrow >= 0:
                rest = rest[:arrow].strip()
            if "[" not in rest:
                if rest not in regs:
                    raise QasmError(f"Invalid qubit token: {rest}")
                off, size = regs[rest]
                targets = list(range(off, off + size))
            else:
                targets = [_qubit_index(regs, rest)]
            for q in targets:
                if q < 0 or q >= n:
                    raise QasmError(f"Target out of bounds: {q}")
            items.append(("measure", targets))
            return
        if s.startswith("bit") or s.startswith("creg") or s.startswith("barrier"):
            return
        if not has_circuit:
            raise QasmError("QASM missing qubit declaration before gates.")
        nctrl = 0
        if s.startswith("ctrl"):
            at = s.find("@")
            if at < 0:
                raise QasmError(f"Unsupported QASM line: {s}")
            mod = s[4:at].strip()
            if not mod:
                nctrl = 1
            elif mod[0] == "(" and mod[-1] == ")":
                nctrl = int(_number(mod[1:-1]))
            else:
                raise QasmError(f"Unsupported QASM line: {s}")
            s = s[at + 1:].strip()
        k = 0
        while k < len(s) and (s[k].isalnum() or s[k] == "_"):
            k += 1
        name, rest = s[:k], s[k:]
        args: List[float] = []
        if rest.strip().startswith("("):
            rest = rest.strip()
            depth, close = 0, -1
            for i, ch in enumerate(rest):
                if ch == "(":
                    depth += 1
                if ch == ")":
                    depth -= 1
                    if depth == 0:
                        close = i
                        break
            if close < 0:
                raise QasmError(f"Invalid {name} gate: {s}")
            args = _split_args(rest[1:close])
            rest = rest[close + 1:]
        q = [_qubit_index(regs, o) for o in _split_operands(rest)]

        def need(nargs, nq):
            if len(args) != nargs:
                raise QasmError(f"Invalid {name} gate: {s}")
            if len(q) != nq + nctrl:
                if name == "cp" and nctrl == 0:
                    raise QasmError(f"Invalid cp operands: {rest.strip()}")
                raise QasmError(f"Invalid {name} operands: {rest.strip()}")

        ctl = q[:nctrl]
        if name in _FIXED:
            need(0, 1)
            items.append(("gate", ctl, q[-1], _FIXED[name]))
        elif name in ("p", "phase", "u1", "s", "sdg", "t", "tdg"):
            if name in ("s", "sdg", "t", "tdg"):
                need(0, 1)
                th = {"s": pi / 2, "sdg": -pi / 2, "t": pi / 4, "tdg": -pi / 4}[name]
            else:
                need(1, 1)
                th = args[0]
            items.append(("gate", ctl, q[-1], P_Gate(th)))
        elif name in _ROT:
            need(1, 1)
            items.append(("gate", ctl, q[-1], _ROT[name](args[0])))
        elif name in ("u3", "U", "u", "u2"):
            if name == "u2":
                need(2, 1)
                g = U3(pi / 2, args[0], args[1])
            else:
                need(3, 1)
                g = U3(args[0], args[1], args[2])
            items.append(("gate", ctl, q[-1], Custom_Gate(g)))
        elif name in ("cx", "CX", "cnot", "cy", "cz"):
            need(0, 2)
            info = Y_Gate if name == "cy" else Z_Gate if name == "cz" else X_Gate
            items.append(("gate", ctl + [q[nctrl]], q[-1], info))
        elif name in ("cp", "cphase", "cu1"):
            need(1, 2)
            items.append(("gate", ctl + [q[nctrl]], q[-1], P_Gate(args[0])))
        elif name in ("crx", "cry", "crz"):
            need(1, 2)
            items.append(("gate", ctl + [q[nctrl]], q[-1], _ROT[name[1:]](args[0])))
        elif name in ("ccx", "toffoli"):
            need(0, 3)
            items.append(("gate", ctl + q[nctrl:nctrl + 2], q[-1], X_Gate))
        elif name == "mcp":
            if len(args) != 1 or len(q) < 2:
                raise QasmError(f"Invalid mcp gate: {s}")
            items.append(("gate", q[:-1], q[-1], P_Gate(args[0])))
        elif name == "swap" and nctrl == 0:
            need(0, 2)
            items.append(("swap", q[0], q[1]))
        else:
            raise QasmError(f"Unsupported QASM line: {s}")

    for line in qasm.split("\n"):
        cm = line.find("//")
        if cm >= 0:
            line = line[:cm]
        for st in line.split(";"):
            statement(st.strip())
    if not has_circuit:
        raise QasmError("No qubit declaration found in QASM.")
    qc = QuantumCircuit(n)
    for it in items:
        if it[0] == "gate":
            _, ctl, t, info = it
            if ctl:
                qc.add(list(ctl), t, info)
            else:
                qc.add(t, info)
        elif it[0] == "swap":
            qc.swap(it[1], it[2])
        else:
            qc.transformations.append(MeasurementTransformation(list(it[1]), None, []))
    return qc


def parse_qasm3_file(path: str) -> QuantumCircuit:
    with open(path, "r", encoding="utf-8") as f:
        return parse_qasm3_string(f.read())
