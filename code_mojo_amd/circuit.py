"""Circuit IR of the Butterfly simulator, mirrored for the GPU strategy.

Transformation kinds, Register offsets, append/c_append and inverse follow
butterfly/core/circuit.mojo:46-961; the classical builders (BITREV, PERMUTE,
SWAP, measure) follow butterfly/core/quantum_circuit.mojo:99-182.  A circuit
is a declarative list; nothing here touches amplitudes.  Lowering to the GPU
op array lives in lowering.py.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple, Union

import numpy as np

from .gates import (GateInfo, GateKind, H_Gate, P_Gate, RX_Gate, RY_Gate, RZ_Gate,
                    X_Gate, Y_Gate, Z_Gate, Gate)


class ControlKind:
    """butterfly/core/circuit.mojo:6-43."""

    NO_CONTROL = 0
    SINGLE_CONTROL = 1
    MULTI_CONTROL = 2


@dataclass
class GateTransformation:
    """controls + target + gate (butterfly/core/circuit.mojo:46-67)."""

    controls: List[int]
    target: int
    gate_info: GateInfo

    @property
    def kind(self) -> int:
        n = len(self.controls)
        return ControlKind.NO_CONTROL if n == 0 else (
            ControlKind.SINGLE_CONTROL if n == 1 else ControlKind.MULTI_CONTROL)


@dataclass
class FusedPairTransformation:
    """Two gates applied together (butterfly/core/circuit.mojo:70-88)."""

    first: GateTransformation
    second: GateTransformation


@dataclass
class SwapTransformation:
    """butterfly/core/circuit.mojo:91-97."""

    a: int
    b: int


@dataclass
class QubitReversalTransformation:
    """butterfly/core/circuit.mojo:100-107 (empty targets = all qubits)."""

    targets: List[int] = field(default_factory=list)


@dataclass
class UnitaryTransformation:
    """Dense 2^m block on [target, target+m), row-major (circuit.mojo:110-138)."""

    u: np.ndarray
    target: int
    m: int
    name: str = "unitary"


@dataclass
class ControlledUnitaryTransformation:
    """butterfly/core/circuit.mojo:141-174."""

    u: np.ndarray
    target: int
    control: int
    m: int
    name: str = "unitary"


@dataclass
class ClassicalTransformation:
    """A named classical state transform (butterfly/core/circuit.mojo:177-193).

    Only the names the GPU strategy understands (BITREV, PERMUTE, SWAP,
    sign_flip_oracle) can be lowered; `apply` is kept for API shape only."""

    name: str
    targets: List[int]
    apply: Optional[Callable] = None
    inverse_apply: Optional[Callable] = None


@dataclass
class MeasurementTransformation:
    """butterfly/core/circuit.mojo:196-227."""

    targets: List[int]
    seed: Optional[int] = None
    values: List[Optional[bool]] = field(default_factory=list)

    def __post_init__(self):
        if not self.values:
            self.values = [None] * len(self.targets)


Transformation = Union[GateTransformation, FusedPairTransformation, SwapTransformation,
                       QubitReversalTransformation, UnitaryTransformation,
                       ControlledUnitaryTransformation, ClassicalTransformation,
                       MeasurementTransformation]


@dataclass
class Register:
    """butterfly/core/circuit.mojo:241-257."""

    name: str
    length: int
    start: int = 0

    def __getitem__(self, idx: int) -> int:
        return self.start + idx


QuantumRegister = Register


def _check_unitary(u, m: int) -> np.ndarray:
    arr = np.asarray(u, dtype=np.complex128).reshape(-1)
    expected = (1 << m) * (1 << m)
    if arr.size != expected:
        raise ValueError(f"Unitary size mismatch. Expected {expected}, got {arr.size}")
    return arr.reshape(1 << m, 1 << m).copy()


class Circuit:
    """Ordered transformation list (butterfly/core/circuit.mojo:259-961)."""

    def __init__(self, arg: Union[int, Register, Sequence[Register]]):
        self.transformations: List[Transformation] = []
        self.registers: List[Register] = []
        if isinstance(arg, int):
            self.num_qubits = arg
        elif isinstance(arg, Register):
            self.registers.append(Register(arg.name, arg.length, arg.start))
            self.num_qubits = arg.length
        else:
            offset = 0
            for reg in arg:
                self.registers.append(Register(reg.name, reg.length, offset))
                offset += reg.length
            self.num_qubits = offset

    # -- generic add (circuit.mojo:286-313) -----------------------------------
    def add(self, *args) -> None:
        """add(target, gate_info) | add(control, target, gate_info) |
        add([controls], target, gate_info)."""
        if len(args) == 2:
            controls, (target, info) = [], args
        elif len(args) == 3:
            c, target, info = args
            controls = list(c) if isinstance(c, (list, tuple)) else [c]
        else:
            raise TypeError("add expects 2 or 3 arguments")
        self.transformations.append(GateTransformation(list(controls), int(target), info))

    def add_classical(self, name: str, targets: Sequence[int], apply=None, inverse_apply=None) -> None:
        self.transformations.append(ClassicalTransformation(name, list(targets), apply, inverse_apply))

    def swap(self, a: int, b: int) -> None:
        self.transformations.append(SwapTransformation(a, b))

    def qubit_reversal(self, targets: Union[None, Sequence[int], Register] = None) -> None:
        if targets is None:
            self.transformations.append(QubitReversalTransformation())
        elif isinstance(targets, Register):
            self.transformations.append(
                QubitReversalTransformation([targets.start + i for i in range(targets.length)]))
        else:
            self.transformations.append(QubitReversalTransformation(list(targets)))

    # -- composition (circuit.mojo:350-662) -----------------------------------
    def append_circuit(self, other: "Circuit", reg: Optional[Register] = None) -> bool:
        """Append `other` with every qubit shifted by reg.start (circuit.mojo:356-501)."""
        if reg is None:
            reg = Register("default", self.num_qubits)
        if other.num_qubits != reg.length:
            return False
        off = reg.start
        for tr in other.transformations:
            self.transformations.append(_shift(tr, off, other.num_qubits))
        return True

    def find_register(self, name: str) -> int:
        for i, r in enumerate(self.registers):
            if r.name == name:
                return i
        return -1

    def append_circuit_by_name(self, name: str, other: "Circuit") -> bool:
        idx = self.find_register(name)
        return False if idx < 0 else self.append_circuit(other, self.registers[idx])

    def c_append_circuit(self, control: int, other: "Circuit", reg: Optional[Register] = None) -> bool:
        """Append `other` controlled by `control` (circuit.mojo:527-662)."""
        if reg is None:
            reg = Register("default", self.num_qubits)
        if other.num_qubits != reg.length:
            return False
        for tr in other.transformations:
            if isinstance(tr, MeasurementTransformation):
                raise ValueError("Controlled measurement is not supported")
            if isinstance(tr, ClassicalTransformation):
                raise ValueError("Controlled classical transformations are not supported")
        off = reg.start

        def cswap(c: int, a: int, b: int) -> None:
            if a == b:
                return
            self.transformations.append(GateTransformation([c, b], a, X_Gate))
            self.transformations.append(GateTransformation([c, a], b, X_Gate))
            self.transformations.append(GateTransformation([c, b], a, X_Gate))

        for tr in other.transformations:
            if isinstance(tr, GateTransformation):
                self.transformations.append(GateTransformation(
                    [control] + [c + off for c in tr.controls], tr.target + off, tr.gate_info))
            elif isinstance(tr, FusedPairTransformation):
                self.transformations.append(FusedPairTransformation(
                    GateTransformation([control] + [c + off for c in tr.first.controls],
                                       tr.first.target + off, tr.first.gate_info),
                    GateTransformation([control] + [c + off for c in tr.second.controls],
                                       tr.second.target + off, tr.second.gate_info)))
            elif isinstance(tr, SwapTransformation):
                cswap(control, tr.a + off, tr.b + off)
            elif isinstance(tr, QubitReversalTransformation):
                targets = sorted(tr.targets) if tr.targets else list(range(other.num_qubits))
                half = len(targets) // 2
                for i in range(half):
                    cswap(control, targets[i] + off, targets[len(targets) - 1 - i] + off)
            elif isinstance(tr, UnitaryTransformation):
                self.transformations.append(ControlledUnitaryTransformation(
                    tr.u.copy(), tr.target + off, control, tr.m, tr.name))
            elif isinstance(tr, ControlledUnitaryTransformation):
                # The reference turns a doubly-controlled unitary into a bare
                # CUSTOM gate and drops the matrix (circuit.mojo:786-808); here
                # the matrix is kept and both controls apply.
                self.transformations.append(_MultiControlledUnitary(
                    tr.u.copy(), tr.target + off, [control, tr.control + off], tr.m, tr.name))
            else:
                raise ValueError("Unsupported transformation in controlled circuit")
        return True

    # -- dense unitaries (circuit.mojo:676-822) -------------------------------
    def unitary(self, u, target: int, m: int = 1, name: str = "unitary") -> None:
        if isinstance(u, np.ndarray) and u.shape == (2, 2) and m == 1:
            u = u.reshape(-1)
        self.transformations.append(UnitaryTransformation(_check_unitary(u, m), target, m, name))

    def u(self, u, target: int) -> None:
        self.unitary(u, target)

    def append_u(self, u, register: Register, name: str = "unitary") -> None:
        self.unitary(u, register.start, register.length, name)

    def c_unitary(self, u, control, target: int, m: int = 1, name: str = "unitary") -> None:
        mat = _check_unitary(u, m)
        if isinstance(control, (list, tuple)):
            self.transformations.append(_MultiControlledUnitary(mat, target, list(control), m, name))
        else:
            self.transformations.append(ControlledUnitaryTransformation(mat, target, control, m, name))

    def cu(self, u, control: int, target: int) -> None:
        self.c_unitary(u, control, target)

    def c_append_u(self, u, control: int, register: Register, name: str = "unitary") -> None:
        self.c_unitary(u, control, register.start, register.length, name)

    # -- gate helpers (circuit.mojo:824-943) ----------------------------------
    def h(self, t: int) -> None: self.add(t, H_Gate)
    def x(self, t: int) -> None: self.add(t, X_Gate)
    def y(self, t: int) -> None: self.add(t, Y_Gate)
    def z(self, t: int) -> None: self.add(t, Z_Gate)
    def p(self, t: int, theta: float) -> None: self.add(t, P_Gate(theta))
    def cp(self, c: int, t: int, theta: float) -> None: self.add(c, t, P_Gate(theta))
    def cx(self, c: int, t: int) -> None: self.add(c, t, X_Gate)
    def ccx(self, c1: int, c2: int, t: int) -> None: self.add([c1, c2], t, X_Gate)
    def cy(self, c: int, t: int) -> None: self.add(c, t, Y_Gate)
    def cz(self, c: int, t: int) -> None: self.add(c, t, Z_Gate)
    def rx(self, t: int, theta: float) -> None: self.add(t, RX_Gate(theta))
    def ry(self, t: int, theta: float) -> None: self.add(t, RY_Gate(theta))
    def rz(self, t: int, theta: float) -> None: self.add(t, RZ_Gate(theta))
    def crx(self, c: int, t: int, theta: float) -> None: self.add(c, t, RX_Gate(theta))
    def cry(self, c: int, t: int, theta: float) -> None: self.add(c, t, RY_Gate(theta))
    def crz(self, c: int, t: int, theta: float) -> None: self.add(c, t, RZ_Gate(theta))
    def mcp(self, controls: Sequence[int], t: int, theta: float) -> None: self.add(list(controls), t, P_Gate(theta))

    def inverse(self) -> "Circuit":
        """Reverse order, invert each transformation (circuit.mojo:945-1045)."""
        inv = Circuit(self.num_qubits)
        for tr in reversed(self.transformations):
            inv.transformations.append(_invert(tr))
        inv.registers = [Register(r.name, r.length, r.start) for r in self.registers]
        return inv


QuantumCircuit = Circuit


@dataclass
class _MultiControlledUnitary:
    """Dense block with several controls (the corrected form of the
    reference's multi-control c_unitary, circuit.mojo:786-808)."""

    u: np.ndarray
    target: int
    controls: List[int]
    m: int
    name: str = "unitary"


def _shift(tr, off: int, other_n: int):
    """Copy of `tr` with qubit indices shifted by `off` (circuit.mojo:364-500)."""
    if isinstance(tr, GateTransformation):
        return GateTransformation([c + off for c in tr.controls], tr.target + off, tr.gate_info)
    if isinstance(tr, FusedPairTransformation):
        return FusedPairTransformation(_shift(tr.first, off, other_n), _shift(tr.second, off, other_n))
    if isinstance(tr, SwapTransformation):
        return SwapTransformation(tr.a + off, tr.b + off)
    if isinstance(tr, QubitReversalTransformation):
        t = tr.targets if tr.targets else list(range(other_n))
        return QubitReversalTransformation([x + off for x in t])
    if isinstance(tr, UnitaryTransformation):
        return UnitaryTransformation(tr.u.copy(), tr.target + off, tr.m, tr.name)
    if isinstance(tr, ControlledUnitaryTransformation):
        return ControlledUnitaryTransformation(tr.u.copy(), tr.target + off, tr.control + off, tr.m, tr.name)
    if isinstance(tr, _MultiControlledUnitary):
        return _MultiControlledUnitary(tr.u.copy(), tr.target + off, [c + off for c in tr.controls], tr.m, tr.name)
    if isinstance(tr, ClassicalTransformation):
        # every classical transform's targets are shifted like qubits, also
        # for index-valued ones such as sign_flip_oracle (circuit.mojo:444-464)
        t = tr.targets if tr.targets else list(range(other_n))
        return ClassicalTransformation(tr.name, [x + off for x in t], tr.apply, tr.inverse_apply)
    if isinstance(tr, MeasurementTransformation):
        return MeasurementTransformation([t + off for t in tr.targets], tr.seed, list(tr.values))
    raise TypeError(f"unknown transformation {type(tr).__name__}")


def _invert_gate(g: GateTransformation) -> GateTransformation:
    """_invert_gate_info (circuit.mojo:1023-1045)."""
    k = g.gate_info.kind
    if k in (GateKind.H, GateKind.X, GateKind.Y, GateKind.Z):
        return GateTransformation(list(g.controls), g.target, g.gate_info)
    if k in (GateKind.P, GateKind.RX, GateKind.RY, GateKind.RZ):
        if g.gate_info.arg is None:
            return GateTransformation(list(g.controls), g.target, g.gate_info)
        return GateTransformation(list(g.controls), g.target, GateInfo.of(k, -g.gate_info.arg))
    raise ValueError("Custom gate inversion not implemented. Use unitary matrices for complex gates.")


def _invert(tr):
    """_invert_transformation (circuit.mojo:964-1014).  Unitary blocks are
    inverted with the true conjugate transpose (the reference only conjugates,
    circuit.mojo:1077-1088)."""
    if isinstance(tr, MeasurementTransformation):
        raise ValueError("Cannot invert circuit containing measurements")
    if isinstance(tr, GateTransformation):
        return _invert_gate(tr)
    if isinstance(tr, FusedPairTransformation):
        return FusedPairTransformation(_invert_gate(tr.second), _invert_gate(tr.first))
    if isinstance(tr, UnitaryTransformation):
        return UnitaryTransformation(tr.u.conj().T.copy(), tr.target, tr.m, tr.name + "_inv")
    if isinstance(tr, ControlledUnitaryTransformation):
        return ControlledUnitaryTransformation(tr.u.conj().T.copy(), tr.target, tr.control, tr.m, tr.name + "_inv")
    if isinstance(tr, _MultiControlledUnitary):
        return _MultiControlledUnitary(tr.u.conj().T.copy(), tr.target, list(tr.controls), tr.m, tr.name + "_inv")
    if isinstance(tr, ClassicalTransformation):
        if tr.inverse_apply is not None:
            return ClassicalTransformation(tr.name + "_inv", list(tr.targets), tr.inverse_apply, tr.apply)
        return tr
    return tr  # swaps and qubit reversal are self-inverse


# -- classical transform builders (quantum_circuit.mojo:99-182) ---------------
def bit_reverse_tr(targets: Optional[Sequence[int]] = None) -> ClassicalTransformation:
    return ClassicalTransformation("BITREV", list(targets) if targets else [])


def permute_qubits_tr(order: Sequence[int]) -> ClassicalTransformation:
    return ClassicalTransformation("PERMUTE", list(order))


def swap_tr(a: int, b: int) -> ClassicalTransformation:
    return ClassicalTransformation("SWAP", [a, b])


def bit_reverse(circuit: Circuit, targets: Optional[Sequence[int]] = None) -> None:
    circuit.transformations.append(bit_reverse_tr(targets))


def permute_qubits(circuit: Circuit, order: Sequence[int]) -> None:
    circuit.transformations.append(permute_qubits_tr(order))


def measure(circuit: Circuit, targets: Sequence[int], values=None, seed: Optional[int] = None) -> None:
    circuit.transformations.append(MeasurementTransformation(list(targets), seed, list(values or [])))


# -- classical replacement rules (quantum_circuit.mojo:33-75, 185-495) ---------
class ClassicalReplacementKind:
    """Which quantum sub-sequences `replace_quantum_with_classical` rewrites
    (butterfly/core/quantum_circuit.mojo:33-75)."""

    SWAP = 0
    QUBIT_REVERSAL = 1
    GROVER_DIFFUSER = 2
    GROVER_INVERSION_0 = 3
    names = {0: "SWAP", 1: "QUBIT_REVERSAL", 2: "GROVER_DIFFUSER", 3: "GROVER_INVERSION_0"}


def _plain(tr, kind: int, target: int) -> bool:
    return (isinstance(tr, GateTransformation) and not tr.controls and tr.gate_info.kind == kind
            and tr.target == target)


def _mcp_pi(tr, base: int, n: int) -> bool:
    """P(pi) on base+n-1 controlled by base .. base+n-2, in that order."""
    return (isinstance(tr, GateTransformation) and tr.gate_info.kind == GateKind.P
            and tr.gate_info.arg is not None and tr.gate_info.arg == math.pi
            and tr.target == base + n - 1 and list(tr.controls) == list(range(base, base + n - 1)))


def _rule_swap(trs, i, out) -> int:
    tr = trs[i]
    if not isinstance(tr, SwapTransformation):
        return 0
    out.append(swap_tr(tr.a, tr.b))
    return 1


def _rule_qrev(trs, i, out) -> int:
    tr = trs[i]
    if not isinstance(tr, QubitReversalTransformation):
        return 0
    out.append(bit_reverse_tr(tr.targets if tr.targets else None))
    return 1


def _run(trs, i, kind) -> Tuple[int, int]:
    """Length and base of the run kind(base), kind(base+1), ... at i."""
    first = trs[i]
    if not (isinstance(first, GateTransformation) and not first.controls and first.gate_info.kind == kind):
        return 0, 0
    base, n = first.target, 0
    while i + n < len(trs) and _plain(trs[i + n], kind, base + n):
        n += 1
    return n, base


def _rule_diffuser(trs, i, out) -> int:
    """H(all) X(all) MCP(pi) X(all) H(all) on base .. base+n-1."""
    n, base = _run(trs, i, GateKind.H)
    if n == 0 or i + 4 * n >= len(trs):
        return 0
    if not all(_plain(trs[i + n + j], GateKind.X, base + j) for j in range(n)):
        return 0
    if not _mcp_pi(trs[i + 2 * n], base, n):
        return 0
    if not all(_plain(trs[i + 2 * n + 1 + j], GateKind.X, base + j) for j in range(n)):
        return 0
    if not all(_plain(trs[i + 3 * n + 1 + j], GateKind.H, base + j) for j in range(n)):
        return 0
    out.append(ClassicalTransformation("GROVER_DIFFUSER", list(range(base, base + n))))
    return 4 * n + 1


def _rule_inversion_0(trs, i, out) -> int:
    """X(all) MCP(pi) X(all) on base .. base+n-1."""
    n, base = _run(trs, i, GateKind.X)
    if n == 0 or i + 2 * n >= len(trs):
        return 0
    if not _mcp_pi(trs[i + n], base, n):
        return 0
    if not all(_plain(trs[i + n + 1 + j], GateKind.X, base + j) for j in range(n)):
        return 0
    out.append(ClassicalTransformation("GROVER_INVERSION_0", list(range(base, base + n))))
    return 2 * n + 1


_RULES = {ClassicalReplacementKind.SWAP: _rule_swap, ClassicalReplacementKind.QUBIT_REVERSAL: _rule_qrev,
          ClassicalReplacementKind.GROVER_DIFFUSER: _rule_diffuser,
          ClassicalReplacementKind.GROVER_INVERSION_0: _rule_inversion_0}


def replace_quantum_with_classical(circuit: Circuit, kinds: Optional[Sequence[int]] = None) -> int:
    """Rewrite matching transformation runs into classical transforms, first
    matching rule wins at each position; returns the number of rewrites
    (apply_classical_rewrites, quantum_circuit.mojo:185-226).  Without
    `kinds` the default rules are SWAP and QUBIT_REVERSAL (:229-233)."""
    if kinds is None:
        kinds = [ClassicalReplacementKind.SWAP, ClassicalReplacementKind.QUBIT_REVERSAL]
    rules = []
    for k in kinds:
        if k not in _RULES:
            raise ValueError(f"Unknown classical replacement kind: {k}")
        rules.append(_RULES[k])
    trs = circuit.transformations
    out: List = []
    i, replaced = 0, 0
    while i < len(trs):
        consumed = 0
        for rule in rules:
            consumed = rule(trs, i, out)
            if consumed > 0:
                replaced += 1
                break
        if consumed <= 0:
            out.append(trs[i])
            i += 1
        else:
            i += consumed
    circuit.transformations = out
    return replaced
