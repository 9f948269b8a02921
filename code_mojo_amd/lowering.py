"""Lowering of a Butterfly circuit into the C-ABI op array (bfly_op).

The GPU strategy crosses the FFI once per circuit with one flat array of
`bfly_op` records (include/bfly_hip.h) plus a side array of dense-block
entries.  The mapping mirrors the reference's per-transformation dispatch in
execute_with_handlers (butterfly/core/executors.mojo:268-319):

  GateTransformation                 -> GATE (kind, target, control mask, 2x2, angle)
  FusedPairTransformation            -> its two gates in order (apply_fused_pair,
                                        transformations_fusion.mojo:221-310, is their product)
  SwapTransformation                 -> SWAP (executors.mojo:231-241: three CX = a bit swap)
  QubitReversalTransformation        -> SWAPs of sorted targets i <-> len-1-i (executors.mojo:244-265)
  Unitary / ControlledUnitary        -> UNITARY (transformations_unitary.mojo:22-139)
  ClassicalTransform "BITREV"        -> BITREV (mask of targets, 0 = all; state.mojo:383-390)
  ClassicalTransform "PERMUTE"       -> the SWAP sequence of apply_permute_qubits (state.mojo:393-431)
  ClassicalTransform "SWAP"          -> SWAP (apply_swap, state.mojo:434-444)
  ClassicalTransform "sign_flip_oracle" -> PHASE_FLIP (mask = all, pattern = index) per listed
                                        index (classical_oracle, algos/grover.mojo:64-66)
  ClassicalTransform "GROVER_INVERSION_0" -> PHASE_FLIP (mask = targets, pattern 0)
                                        (apply_grover_inversion_0, algos/grover_classical.mojo:262-281)
  ClassicalTransform "GROVER_DIFFUSER" -> DIFFUSE (mask = targets, reference variant)
                                        (apply_grover_diffuser, algos/grover_classical.mojo:284-326)
  MeasurementTransformation          -> MEASURE ((qubit, value) pairs in udata, seed; apply_measure,
                                        state.mojo:447-553)
  other classical                    -> error: not on the GPU path

Validation follows validate_circuit (butterfly/core/executors.mojo:90-185).
"""
from __future__ import annotations

from typing import List, Optional, Tuple

import numpy as np

from . import circuit as C
from .gates import GateKind, X, Z

OP_DTYPE = np.dtype([
    ("kind", "<i4"), ("gate", "<i4"), ("target", "<i4"), ("arg_i", "<i4"),
    ("controls", "<u8"), ("aux", "<u8"), ("theta", "<f8"), ("mat", "<f8", (8,)),
], align=True)
assert OP_DTYPE.itemsize == 104

OP_GATE, OP_BITREV, OP_SWAP, OP_UNITARY, OP_PHASE_FLIP, OP_DIFFUSE, OP_MEASURE = 0, 1, 2, 3, 4, 5, 6


class LoweringError(ValueError):
    """A transformation cannot be lowered to the GPU op array."""


def _mat8(g: np.ndarray) -> List[float]:
    g = np.asarray(g, dtype=np.complex128).reshape(2, 2)
    out = []
    for r in range(2):
        for c in range(2):
            out += [float(g[r, c].real), float(g[r, c].imag)]
    return out


class OpList:
    """Accumulates lowered records and dense-block data."""

    def __init__(self, num_qubits: int):
        self.n = num_qubits
        self.rows: List[tuple] = []
        self.udata: List[np.ndarray] = []
        self.ulen = 0

    def gate(self, target: int, controls: List[int], kind: int, g: np.ndarray, theta: float = 0.0):
        mask = 0
        for c in controls:
            mask |= 1 << int(c)
        self.rows.append((OP_GATE, int(kind), int(target), 0, mask, 0, float(theta), _mat8(g)))

    def bitrev(self, mask: int):
        self.rows.append((OP_BITREV, 0, 0, 0, 0, int(mask), 0.0, [0.0] * 8))

    def swap(self, a: int, b: int):
        self.rows.append((OP_SWAP, 0, int(a), int(b), 0, 0, 0.0, [0.0] * 8))

    def phase_flip(self, mask: int, value: int):
        self.rows.append((OP_PHASE_FLIP, 0, 0, 0, int(value), int(mask), 0.0, [0.0] * 8))

    def diffuse(self, mask: int, variant: int):
        self.rows.append((OP_DIFFUSE, 0, 0, int(variant), 0, int(mask), 0.0, [0.0] * 8))

    def measure(self, targets: List[int], values: List[Optional[bool]], seed: Optional[int]):
        """MEASURE record: (qubit, value) pairs in udata, value -1 = free."""
        inter = np.empty(2 * len(targets), dtype=np.float64)
        for k, (t, v) in enumerate(zip(targets, values)):
            inter[2 * k] = int(t)
            inter[2 * k + 1] = -1 if v is None else (1 if v else 0)
        has_seed = seed is not None and seed >= 0
        self.rows.append((OP_MEASURE, 1 if has_seed else 0, 0, len(targets), int(seed) if has_seed else 0,
                          self.ulen, 0.0, [0.0] * 8))
        self.udata.append(inter)
        self.ulen += len(targets)

    def unitary(self, u: np.ndarray, target: int, m: int, controls: List[int]):
        mask = 0
        for c in controls:
            mask |= 1 << int(c)
        flat = np.asarray(u, dtype=np.complex128).reshape(-1)
        self.rows.append((OP_UNITARY, 0, int(target), int(m), mask, self.ulen, 0.0, [0.0] * 8))
        inter = np.empty(2 * flat.size, dtype=np.float64)
        inter[0::2] = flat.real
        inter[1::2] = flat.imag
        self.udata.append(inter)
        self.ulen += flat.size

    def finish(self) -> Tuple[np.ndarray, np.ndarray]:
        ops = np.zeros(len(self.rows), dtype=OP_DTYPE)
        for i, r in enumerate(self.rows):
            ops[i] = r
        udata = np.concatenate(self.udata) if self.udata else np.zeros(0, dtype=np.float64)
        return ops, udata


def _check_q(q: int, n: int, label: str) -> None:
    if q < 0 or q >= n:
        raise LoweringError(f"{label} out of bounds: {q}")


def _lower_gate(out: OpList, g: C.GateTransformation) -> None:
    n = out.n
    _check_q(g.target, n, "Target")
    for c in g.controls:
        _check_q(c, n, "Control")
    if g.target in g.controls:
        # validate_gate_transformation (executors.mojo:95-108) accepts this,
        # and the scalar kernels test the control on the pair's low index,
        # whose target bit is 0 (transformations_scalar.mojo:160,590): the
        # gate never acts.
        return
    theta = g.gate_info.arg if g.gate_info.arg is not None else 0.0
    out.gate(g.target, g.controls, g.gate_info.kind, g.gate_info.gate, theta)


def _lower_permute(out: OpList, order: List[int], inverse: bool = False) -> None:
    """apply_permute_qubits (state.mojo:393-431): selection of swaps; with
    `inverse` the same swaps in reverse order (the inverse permutation)."""
    n = len(order)
    if n == 0:
        return
    if n != out.n:
        raise LoweringError(f"Permutation size mismatch: {n}")
    seen = [False] * n
    for v in order:
        if v < 0 or v >= n:
            raise LoweringError(f"Invalid permutation value: {v}")
        if seen[v]:
            raise LoweringError(f"Duplicate permutation value: {v}")
        seen[v] = True
    perm = list(range(n))
    swaps = []
    for i in range(n):
        if perm[i] == order[i]:
            continue
        j = i + 1
        while j < n and perm[j] != order[i]:
            j += 1
        if j >= n:
            raise LoweringError("Invalid permutation order")
        swaps.append((i, j))
        perm[i], perm[j] = perm[j], perm[i]
    for a, b in (reversed(swaps) if inverse else swaps):
        out.swap(a, b)


def _lower_sign_flip(out: OpList, indices: List[int]) -> None:
    n = out.n
    for m in indices:
        if m < 0 or m >= (1 << n):
            raise LoweringError(f"Target out of bounds: {m}")
        out.phase_flip((1 << n) - 1, m)


def _target_mask(targets: List[int], n: int) -> int:
    """_normalize_targets (algos/grover_classical.mojo:221-245): empty means
    every qubit; out of range and duplicates are errors."""
    if not targets:
        return (1 << n) - 1
    mask = 0
    for t in targets:
        if t < 0 or t >= n:
            raise LoweringError(f"Target out of bounds: {t}")
        if (mask >> t) & 1:
            raise LoweringError(f"Duplicate target: {t}")
        mask |= 1 << t
    return mask


def _lower_measure(out: OpList, tr, n: int, measure_seed: Optional[int]) -> None:
    """apply_measure's checks (state.mojo:447-462): values either empty (all
    free) or one per target; targets within the register.  A transformation
    without a seed takes `measure_seed` (the reference's configured
    measure_seed), else the clock at execution time."""
    targets = [int(t) for t in tr.targets]
    if not targets:
        return
    values = list(tr.values) if tr.values else [None] * len(targets)
    if len(values) != len(targets):
        raise LoweringError("Measurement values size mismatch")
    for t in targets:
        if t < 0 or t >= n:
            raise LoweringError(f"Target out of bounds: {t}")
    seed = tr.seed if tr.seed is not None else measure_seed
    out.measure(targets, values, seed)


def lower_circuit(circuit: C.Circuit, num_qubits: int | None = None,
                  measure_seed: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Return (ops, udata) for `circuit`; raises LoweringError on anything the
    GPU strategy does not run (validate_circuit semantics)."""
    n = circuit.num_qubits if num_qubits is None else num_qubits
    if circuit.num_qubits != n:
        raise LoweringError(
            f"Circuit qubits mismatch: circuit has {circuit.num_qubits}, state has {n}")
    out = OpList(n)
    for tr in circuit.transformations:
        if isinstance(tr, C.GateTransformation):
            _lower_gate(out, tr)
        elif isinstance(tr, C.FusedPairTransformation):
            _lower_gate(out, tr.first)
            _lower_gate(out, tr.second)
        elif isinstance(tr, C.SwapTransformation):
            _check_q(tr.a, n, "Target")
            _check_q(tr.b, n, "Target")
            if tr.a != tr.b:
                out.swap(tr.a, tr.b)
        elif isinstance(tr, C.QubitReversalTransformation):
            t = sorted(tr.targets) if tr.targets else list(range(n))
            for q in t:
                _check_q(q, n, "Target")
            for i in range(len(t) // 2):
                a, b = t[i], t[len(t) - 1 - i]
                if a != b:
                    out.swap(a, b)
        elif isinstance(tr, C.UnitaryTransformation):
            if tr.m <= 0:
                continue
            if tr.target < 0 or tr.target + tr.m > n:
                raise LoweringError("Target range out of bounds")
            out.unitary(tr.u, tr.target, tr.m, [])
        elif isinstance(tr, (C.ControlledUnitaryTransformation, C._MultiControlledUnitary)):
            if tr.m <= 0:
                continue
            if tr.target < 0 or tr.target + tr.m > n:
                raise LoweringError("Target range out of bounds")
            ctrls = [tr.control] if isinstance(tr, C.ControlledUnitaryTransformation) else list(tr.controls)
            for c in ctrls:
                _check_q(c, n, "Control")
                if tr.target <= c < tr.target + tr.m:
                    raise LoweringError("Control overlaps target range")
            out.unitary(tr.u, tr.target, tr.m, ctrls)
        elif isinstance(tr, C.ClassicalTransformation):
            # NAME_inv: the reference renames a classical transform on inversion
            # when it carries an inverse_apply (circuit.mojo:1000-1010); BITREV,
            # SWAP, sign flips and inversion-0 are their own inverses, PERMUTE's
            # inverse is its swap sequence reversed
            inverse = tr.name.endswith("_inv")
            name = tr.name[:-4] if inverse else tr.name
            if inverse and name == "GROVER_DIFFUSER":
                raise LoweringError("GROVER_DIFFUSER_inv: the classical diffuser (im <- -2 mean_im - im) is not "
                                    "its own inverse and has no GPU lowering")
            if name == "BITREV":
                mask = 0
                for q in tr.targets:
                    _check_q(q, n, "Target")
                    mask |= 1 << q
                if len(tr.targets) == 0:
                    out.bitrev(0)
                elif len(tr.targets) > 1:
                    out.bitrev(mask)
            elif name == "PERMUTE":
                _lower_permute(out, list(tr.targets), inverse)
            elif name == "SWAP":
                if len(tr.targets) != 2:
                    raise LoweringError("SWAP expects 2 targets")
                a, b = tr.targets
                _check_q(a, n, "Target")
                _check_q(b, n, "Target")
                if a != b:
                    out.swap(a, b)
            elif name == "sign_flip_oracle":
                _lower_sign_flip(out, list(tr.targets))
            elif name == "GROVER_INVERSION_0":
                out.phase_flip(_target_mask(list(tr.targets), n), 0)
            elif name == "GROVER_DIFFUSER":
                out.diffuse(_target_mask(list(tr.targets), n), 1)
            else:
                raise LoweringError(
                    f"classical transform '{tr.name}' cannot be lowered to the GPU strategy")
        elif isinstance(tr, C.MeasurementTransformation):
            _lower_measure(out, tr, n, measure_seed)
        else:
            raise LoweringError("Unknown transformation type")
    return out.finish()
