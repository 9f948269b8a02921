"""Host-side fusion rewrites of a circuit's transformation list
(butterfly/core/transformations_fusion.mojo:24-908), restated for the GPU
strategy's Python mirror.

The reference's fused strategies run circuits rewritten by these passes; they
emit the same transformation types the GPU strategy consumes:
  fuse_circuit                     contiguous-target pairs of uncontrolled gates
                                   -> UnitaryTransformation(m = 2, "fused_pair"),
                                   then runs of gates on one target -> one P
                                   (summed angle), UnitaryTransformation(m = 1,
                                   "fused") or ControlledUnitaryTransformation
                                   (m = 1, "fused_ctrl")            :879-900
  fuse_circuit_specialized_pairs   HH / HP / PP / CP.CP (shared control) pairs
                                   -> FusedPairTransformation, then the same
                                   same-target pass                   :886-908
4x4 matrices use the local index b0 + 2 b1 with b0 the lower qubit
(gate_to_matrix4x4 / controlled_gate_to_matrix4x4, :51-93), the index
convention of apply_unitary on [low, low + 2).
"""
from __future__ import annotations

from typing import List, Optional, Tuple

import numpy as np

from . import gates as G
from .circuit import (Circuit, ControlKind, ControlledUnitaryTransformation, FusedPairTransformation,
                      GateTransformation, UnitaryTransformation)


# -- matrix helpers (transformations_fusion.mojo:24-117) -----------------------
def gate_to_matrix4x4(gate: np.ndarray, q_low: int, target: int) -> np.ndarray:
    """A 2x2 gate on `target` (q_low or q_low + 1) as a 4x4 matrix."""
    return np.kron(np.eye(2), gate) if target == q_low else np.kron(gate, np.eye(2))


def controlled_gate_to_matrix4x4(gate: np.ndarray, q_low: int, control: int, target: int) -> np.ndarray:
    """A singly controlled 2x2 gate on the pair (q_low, q_low + 1) as a 4x4."""
    m = np.zeros((4, 4), dtype=np.complex128)
    for b0 in range(2):
        for b1 in range(2):
            col = b0 + 2 * b1
            ctrl_bit = b0 if control == q_low else b1
            tgt_bit = b0 if target == q_low else b1
            if ctrl_bit == 0:
                m[col, col] = 1
                continue
            for out_t in range(2):
                o0, o1 = (out_t, b1) if target == q_low else (b0, out_t)
                m[o0 + 2 * o1, col] = gate[out_t, tgt_bit]
    return m


def _bounds(t: GateTransformation) -> Tuple[int, int]:
    qs = [t.target] + (list(t.controls) if t.kind == ControlKind.SINGLE_CONTROL else [])
    return min(qs), max(qs)


def _is_h(t: GateTransformation) -> bool:
    return t.kind == ControlKind.NO_CONTROL and t.gate_info.kind == G.GateKind.H


def _is_p(t: GateTransformation) -> bool:
    return t.kind == ControlKind.NO_CONTROL and t.gate_info.kind == G.GateKind.P


def _is_cp(t: GateTransformation) -> bool:
    return t.kind == ControlKind.SINGLE_CONTROL and t.gate_info.kind == G.GateKind.P


def is_special_pair_gate(t: GateTransformation) -> bool:
    """H or P without controls, or P with one control (:117-128)."""
    if t.kind == ControlKind.NO_CONTROL:
        return t.gate_info.kind in (G.GateKind.H, G.GateKind.P)
    return t.kind == ControlKind.SINGLE_CONTROL and t.gate_info.kind == G.GateKind.P


def can_fuse_special_pair(t0: GateTransformation, t1: GateTransformation) -> bool:
    """The pairs the sparse fused kernels handle (:181-219)."""
    if t0.kind == ControlKind.MULTI_CONTROL or t1.kind == ControlKind.MULTI_CONTROL:
        return False
    if _is_cp(t0) and _is_cp(t1) and t0.controls[0] == t1.controls[0] and t0.target != t1.target:
        return True
    lo0, hi0 = _bounds(t0)
    lo1, hi1 = _bounds(t1)
    if max(hi0, hi1) - min(lo0, lo1) != 1:
        return False
    return ((_is_h(t0) and _is_h(t1)) or (_is_p(t0) and _is_p(t1)) or (_is_h(t0) and _is_cp(t1))
            or (_is_cp(t0) and _is_h(t1)))


# -- passes --------------------------------------------------------------------
def fuse_special_pair_gates(trs: list) -> list:
    """Consecutive fusable gate pairs -> FusedPairTransformation (:312-333)."""
    out, i = [], 0
    while i < len(trs):
        tr = trs[i]
        if (i + 1 < len(trs) and isinstance(tr, GateTransformation) and isinstance(trs[i + 1], GateTransformation)
                and can_fuse_special_pair(tr, trs[i + 1])):
            out.append(FusedPairTransformation(tr, trs[i + 1]))
            i += 2
            continue
        out.append(tr)
        i += 1
    return out


class _SameTarget:
    """The run of gates on one (target, control) being merged (:334-384)."""

    def __init__(self, t: GateTransformation):
        self.target = t.target
        self.kind = t.kind
        self.control = t.controls[0] if t.kind == ControlKind.SINGLE_CONTROL else 0
        self.gate = np.array(t.gate_info.gate, dtype=np.complex128)
        self.info = t.gate_info
        self.count = 1
        self.is_phase = t.gate_info.kind == G.GateKind.P and t.gate_info.arg is not None
        self.phase_sum = float(t.gate_info.arg) if self.is_phase else 0.0

    def accepts(self, t: GateTransformation) -> bool:
        control = t.controls[0] if t.kind == ControlKind.SINGLE_CONTROL else 0
        return t.target == self.target and t.kind == self.kind and control == self.control

    def add(self, t: GateTransformation) -> None:
        self.gate = t.gate_info.gate @ self.gate
        self.count += 1
        if self.is_phase and t.gate_info.kind == G.GateKind.P and t.gate_info.arg is not None:
            self.phase_sum += float(t.gate_info.arg)
        else:
            self.is_phase = False

    def flush(self, out: list) -> None:
        controls = [self.control] if self.kind == ControlKind.SINGLE_CONTROL else []
        if self.is_phase:
            out.append(GateTransformation(controls, self.target, G.P_Gate(self.phase_sum)))
        elif self.count == 1:
            out.append(GateTransformation(controls, self.target, self.info))
        elif self.kind == ControlKind.NO_CONTROL:
            out.append(UnitaryTransformation(self.gate.copy(), self.target, 1, "fused"))
        else:
            out.append(ControlledUnitaryTransformation(self.gate.copy(), self.target, self.control, 1, "fused_ctrl"))


def fuse_same_target_gates(trs: list) -> list:
    """Runs of gates with the same target and control -> one transformation
    (:427-537)."""
    out: list = []
    active: Optional[_SameTarget] = None
    for tr in trs:
        if isinstance(tr, GateTransformation) and tr.kind != ControlKind.MULTI_CONTROL:
            if active is not None and active.accepts(tr):
                active.add(tr)
                continue
            if active is not None:
                active.flush(out)
            active = _SameTarget(tr)
            continue
        if active is not None:
            active.flush(out)
            active = None
        out.append(tr)
    if active is not None:
        active.flush(out)
    return out


def fuse_contiguous_target_pairs(trs: list) -> list:
    """Uncontrolled gates on two adjacent qubits -> UnitaryTransformation(m = 2)
    (:539-660).  A pending single gate flushes unchanged."""
    out: list = []
    mode = 0            # 0 idle, 1 one pending gate, 2 an active pair
    pending: Optional[GateTransformation] = None
    pair_low = pair_high = 0
    matrix = np.eye(4, dtype=np.complex128)

    def flush():
        nonlocal mode
        if mode == 1:
            out.append(pending)
        elif mode == 2:
            out.append(UnitaryTransformation(matrix.copy(), pair_low, 2, "fused_pair"))
        mode = 0

    for tr in trs:
        if not isinstance(tr, GateTransformation) or tr.kind != ControlKind.NO_CONTROL:
            flush()
            out.append(tr)
            continue
        t = tr.target
        if mode == 0:
            mode, pending = 1, tr
        elif mode == 1:
            if t != pending.target and abs(t - pending.target) == 1:
                pair_low, pair_high = min(t, pending.target), max(t, pending.target)
                matrix = (gate_to_matrix4x4(tr.gate_info.gate, pair_low, t)
                          @ gate_to_matrix4x4(pending.gate_info.gate, pair_low, pending.target))
                mode = 2
            else:
                flush()
                mode, pending = 1, tr
        else:
            if t in (pair_low, pair_high):
                matrix = gate_to_matrix4x4(tr.gate_info.gate, pair_low, t) @ matrix
            else:
                flush()
                mode, pending = 1, tr
    flush()
    return out


def fuse_transformations(trs: list) -> list:
    """fuse_contiguous_target_pairs, then fuse_same_target_gates (:879-884)."""
    return fuse_same_target_gates(fuse_contiguous_target_pairs(trs))


def fuse_transformations_specialized_pairs(trs: list) -> list:
    """fuse_special_pair_gates, then fuse_same_target_gates (:886-890)."""
    return fuse_same_target_gates(fuse_special_pair_gates(trs))


def _rewritten(circuit: Circuit, trs: list) -> Circuit:
    out = type(circuit)(circuit.num_qubits)
    out.registers = list(circuit.registers)
    out.transformations = trs
    return out


def fuse_circuit(circuit: Circuit) -> Circuit:
    """butterfly/core/transformations_fusion.mojo:892-898."""
    return _rewritten(circuit, fuse_transformations(list(circuit.transformations)))


def fuse_circuit_specialized_pairs(circuit: Circuit) -> Circuit:
    """butterfly/core/transformations_fusion.mojo:900-908."""
    return _rewritten(circuit, fuse_transformations_specialized_pairs(list(circuit.transformations)))
