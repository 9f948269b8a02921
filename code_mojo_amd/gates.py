"""Gate library of the Butterfly simulator, mirrored for the GPU strategy.

Matrices follow butterfly/core/gates.mojo:5-61 (Qiskit U3 convention) and the
GateKind numbering of butterfly/core/gates.mojo:63-135.  Gates are 2x2 complex
numpy arrays computed in FP64 on the host; the GPU receives them through the
lowered op array.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

Gate = np.ndarray  # shape (2, 2), complex128

pi = math.pi
SQ_HALF = math.sqrt(0.5)  # sq_half_re, butterfly/core/types.mojo:23


def cis(theta: float) -> complex:
    """cos(theta) + i sin(theta) (butterfly/utils/common.mojo:8-9)."""
    return complex(math.cos(theta), math.sin(theta))


def _g(a, b, c, d) -> Gate:
    return np.array([[a, b], [c, d]], dtype=np.complex128)


X: Gate = _g(0, 1, 1, 0)
Y: Gate = _g(0, -1j, 1j, 0)
Z: Gate = _g(1, 0, 0, -1)
H: Gate = _g(SQ_HALF, SQ_HALF, SQ_HALF, -SQ_HALF)


def P(theta: float) -> Gate:
    """diag(1, e^{i theta}) (butterfly/core/gates.mojo:10-12)."""
    return _g(1, 0, 0, cis(theta))


def RZ(theta: float) -> Gate:
    """butterfly/core/gates.mojo:18-23."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return _g(complex(c, -s), 0, 0, complex(c, s))


def RX(theta: float) -> Gate:
    """butterfly/core/gates.mojo:26-31."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return _g(c, complex(0, -s), complex(0, -s), c)


def RY(theta: float) -> Gate:
    """butterfly/core/gates.mojo:34-39."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return _g(c, -s, s, c)


def U3(theta: float, phi: float, lam: float) -> Gate:
    """butterfly/core/gates.mojo:42-53."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return _g(c, -cis(lam) * s, cis(phi) * s, cis(phi + lam) * c)


def U2(phi: float, lam: float) -> Gate:
    return U3(pi / 2, phi, lam)


def U1(lam: float) -> Gate:
    return U3(0.0, 0.0, lam)


class GateKind:
    """Gate kinds with the reference's integer values (gates.mojo:83-91)."""

    H = 0
    P = 1
    X = 2
    Y = 3
    Z = 4
    RX = 5
    RY = 6
    RZ = 7
    CUSTOM = 8

    names = ["H", "P", "X", "Y", "Z", "RX", "RY", "RZ", "CUSTOM"]

    @staticmethod
    def from_int(value: int) -> int:
        if not 0 <= value <= GateKind.CUSTOM:
            raise ValueError(f"Unknown gate kind: {value}")
        return value


_PARAMETRIC = {GateKind.P: P, GateKind.RX: RX, GateKind.RY: RY, GateKind.RZ: RZ}
# gate_types table: CUSTOM's default entry is H (gates.mojo:121-131).
_FIXED = {GateKind.H: H, GateKind.X: X, GateKind.Y: Y, GateKind.Z: Z, GateKind.CUSTOM: H}
DIAGONAL_KINDS = frozenset({GateKind.P, GateKind.Z, GateKind.RZ})


@dataclass(frozen=True)
class GateInfo:
    """Gate kind + 2x2 matrix + optional angle (butterfly/core/gates.mojo:137-175)."""

    kind: int
    gate: Gate
    arg: Optional[float] = None
    name: str = ""

    @staticmethod
    def of(kind: int, arg: Optional[float] = None) -> "GateInfo":
        kind = GateKind.from_int(kind)
        if kind in _PARAMETRIC:
            if arg is None:
                raise ValueError(f"gate {GateKind.names[kind]} needs an angle")
            return GateInfo(kind, _PARAMETRIC[kind](float(arg)), float(arg), GateKind.names[kind])
        return GateInfo(kind, _FIXED[kind], None, GateKind.names[kind])

    def __str__(self) -> str:
        return self.name + (f"({self.arg})" if self.arg is not None else "")


H_Gate = GateInfo.of(GateKind.H)
X_Gate = GateInfo.of(GateKind.X)
Y_Gate = GateInfo.of(GateKind.Y)
Z_Gate = GateInfo.of(GateKind.Z)


def P_Gate(theta: float) -> GateInfo:
    return GateInfo.of(GateKind.P, theta)


def RX_Gate(theta: float) -> GateInfo:
    return GateInfo.of(GateKind.RX, theta)


def RY_Gate(theta: float) -> GateInfo:
    return GateInfo.of(GateKind.RY, theta)


def RZ_Gate(theta: float) -> GateInfo:
    return GateInfo.of(GateKind.RZ, theta)


def Custom_Gate(gate: Gate, name: str = "CUSTOM") -> GateInfo:
    """A custom 2x2 gate, e.g. Custom_Gate(U3(...)) (butterfly/core/gates.mojo:194-195)."""
    g = np.asarray(gate, dtype=np.complex128)
    if g.shape != (2, 2):
        raise ValueError("custom gate must be 2x2")
    return GateInfo(GateKind.CUSTOM, g.copy(), None, name)


def is_diagonal(info: GateInfo) -> bool:
    """True when the gate's matrix is diagonal (phase-only)."""
    if info.kind in DIAGONAL_KINDS:
        return True
    g = info.gate
    return g[0, 1] == 0 and g[1, 0] == 0
