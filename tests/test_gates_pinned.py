"""Gate matrices pinned independently of code_mojo_amd/gates.py.

The reference defines RX / RY / RZ / U3 / Y (butterfly/core/gates.mojo:5-61)
as closed forms.  Here each is checked against mathematics that shares no code
with it: matrix exponentials of the Pauli generators (scipy.linalg.expm),
U3(theta, phi, lambda) = e^{i (phi + lambda) / 2} RZ(phi) RY(theta) RZ(lambda),
and the special cases U3(theta, -pi/2, pi/2) = RX(theta), U3(theta, 0, 0) =
RY(theta), U1 = P, P(theta) = e^{i theta / 2} RZ(theta), Y = i X Z, H = (X + Z)
/ sqrt 2.  The matrices are checked as the lowering hands them to the GPU (the
mat field of the op records), not only as gates.py returns them.  A planted
sign error (RX with +i sin) fails the exponential check."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from code_mojo_amd import gates as G
from code_mojo_amd import lower_circuit
from code_mojo_amd.circuit import QuantumCircuit

# Pauli matrices written out here (not taken from gates.py)
SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, -1j], [1j, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2, dtype=complex)

ANGLES = [0.0, 0.3, -1.1, math.pi / 2, math.pi, 2.5, -3.0, 2 * math.pi, 7.7]
TOL = 4e-16 * 8


def rot(pauli, theta):
    return expm(-0.5j * theta * pauli)


def lowered_matrix(add):
    """The 2x2 matrix of the single gate `add(qc)` puts on qubit 0, as lowered."""
    qc = QuantumCircuit(1)
    add(qc)
    ops, _ = lower_circuit(qc)
    assert ops.size == 1
    m = ops[0]["mat"]
    return np.array([[m[0] + 1j * m[1], m[2] + 1j * m[3]], [m[4] + 1j * m[5], m[6] + 1j * m[7]]])


@pytest.mark.parametrize("theta", ANGLES)
def test_rotations_are_exponentials_of_pauli_generators(theta):
    assert np.abs(G.RX(theta) - rot(SX, theta)).max() <= TOL
    assert np.abs(G.RY(theta) - rot(SY, theta)).max() <= TOL
    assert np.abs(G.RZ(theta) - rot(SZ, theta)).max() <= TOL
    # as lowered for the GPU (GateInfo -> bfly_op.mat)
    assert np.abs(lowered_matrix(lambda q: q.rx(0, theta)) - rot(SX, theta)).max() <= TOL
    assert np.abs(lowered_matrix(lambda q: q.ry(0, theta)) - rot(SY, theta)).max() <= TOL
    assert np.abs(lowered_matrix(lambda q: q.rz(0, theta)) - rot(SZ, theta)).max() <= TOL


@pytest.mark.parametrize("theta,phi,lam", [(0.3, 1.2, -0.7), (2.9, -2.0, 0.4), (math.pi, 0.5, 0.5),
                                           (0.0, 1.0, 2.0), (1.7, 6.0, -6.0)])
def test_u3_is_the_zyz_euler_product(theta, phi, lam):
    ref = np.exp(0.5j * (phi + lam)) * rot(SZ, phi) @ rot(SY, theta) @ rot(SZ, lam)
    assert np.abs(G.U3(theta, phi, lam) - ref).max() <= TOL
    got = lowered_matrix(lambda q: q.add(0, G.Custom_Gate(G.U3(theta, phi, lam))))
    assert np.abs(got - ref).max() <= TOL


@pytest.mark.parametrize("theta", ANGLES)
def test_special_cases(theta):
    assert np.abs(G.U3(theta, -math.pi / 2, math.pi / 2) - G.RX(theta)).max() <= TOL
    assert np.abs(G.U3(theta, 0.0, 0.0) - G.RY(theta)).max() <= TOL
    assert np.abs(G.U1(theta) - G.P(theta)).max() <= TOL
    assert np.abs(G.P(theta) - np.exp(0.5j * theta) * rot(SZ, theta)).max() <= TOL
    u2 = np.array([[1, -np.exp(1j * theta)], [np.exp(0.4j), np.exp(1j * (0.4 + theta))]]) / math.sqrt(2)
    assert np.abs(G.U2(0.4, theta) - u2).max() <= TOL


def test_fixed_gates():
    assert np.array_equal(G.X, SX) and np.array_equal(G.Z, SZ)
    assert np.abs(G.Y - 1j * SX @ SZ).max() == 0.0
    assert np.abs(G.H - (SX + SZ) / math.sqrt(2)).max() <= 2e-16   # sqrt(0.5) vs 1/sqrt(2): one ulp
    for g in (G.X, G.Y, G.Z, G.H):
        assert np.abs(g @ g.conj().T - I2).max() <= 1e-15
    assert np.abs(lowered_matrix(lambda q: q.y(0)) - SY).max() == 0.0


def test_a_planted_sign_error_fails_the_exponential_check():
    def rx_wrong(theta):
        c, s = math.cos(theta / 2), math.sin(theta / 2)
        return np.array([[c, 1j * s], [1j * s, c]])
    assert np.abs(rx_wrong(0.3) - rot(SX, 0.3)).max() > 1e-3
