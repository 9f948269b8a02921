"""CPU tests of the fusion rewrites (code_mojo_amd/fusion.py, restating
butterfly/core/transformations_fusion.mojo:24-908) and of the planner's
handling of the blocks they emit.

Pins: a fused circuit is the same unitary as the original (the scalar oracle
runs both; the oracle's apply_unitary is itself pinned by the Kronecker-product
fixtures), the emitted types follow the reference's rules (adjacent-qubit
pairs, same-target runs, special pairs, merged P angles), the 4x4 helpers
match explicit Kronecker products, and the compiled schedule of a fused
circuit (m = 1 blocks and factorable m = 2 blocks as gate records) replays to
the oracle's amplitudes with no stand-alone unitary launches."""
import math

import numpy as np
import pytest

from code_mojo_amd import _lib, algos, fusion, lower_circuit
from code_mojo_amd import gates as G
from code_mojo_amd import workloads as W
from code_mojo_amd.circuit import (ControlledUnitaryTransformation, FusedPairTransformation, GateTransformation,
                                   QuantumCircuit, UnitaryTransformation)
from code_mojo_amd.state import Plan
from oracle import oracle as O

from .plan_sim import run_plan


def oracle_run(n, qc, psi):
    ops, ud = lower_circuit(qc)
    re, im = psi.real.copy(), psi.imag.copy()
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


def circuits(n):
    qc = QuantumCircuit(n)          # a mix with every pattern the passes look for
    qc.h(0); qc.h(1); qc.p(2, 0.3); qc.p(2, 0.4); qc.cp(0, 1, 0.7); qc.cp(0, 2, 0.2)
    qc.h(3); qc.p(2, -1.1); qc.x(1); qc.ry(1, 0.8); qc.rz(1, 0.2); qc.cx(2, 3); qc.mcp([0, 1], 3, 0.9)
    qc.h(2); qc.h(3); qc.h(2); qc.rx(0, 1.3); qc.p(1, 0.5); qc.cp(3, 0, 0.6); qc.cp(3, 0, 0.1)
    return {
        "mix": qc,
        "qft": W.qft_circuit(n),
        "value": algos.encode_value_circuit(n, 45.0),
        "layered": W.random_layered(n, 4, 5),
    }


@pytest.mark.parametrize("which", ["mix", "qft", "value", "layered"])
@pytest.mark.parametrize("pass_", [fusion.fuse_circuit, fusion.fuse_circuit_specialized_pairs])
def test_fused_circuit_is_the_same_unitary(which, pass_):
    n = 7
    qc = circuits(n)[which]
    fc = pass_(qc)
    assert len(fc.transformations) <= len(qc.transformations)
    psi = W.normal_state(n, 3).to_complex()
    assert np.abs(oracle_run(n, fc, psi) - oracle_run(n, qc, psi)).max() <= 1e-13


def test_emitted_types_follow_the_reference_rules():
    n = 4
    qc = QuantumCircuit(n)
    qc.h(0); qc.h(1)                 # adjacent uncontrolled -> 4x4 pair
    qc.p(3, 0.25); qc.p(3, 0.5)      # same target P run -> one P, summed angle
    fc = fusion.fuse_circuit(qc)
    kinds = [type(t).__name__ for t in fc.transformations]
    assert kinds == ["UnitaryTransformation", "GateTransformation"]
    pair = fc.transformations[0]
    assert pair.m == 2 and pair.target == 0 and pair.name == "fused_pair"
    assert np.abs(pair.u - np.kron(G.H, G.H)).max() <= 1e-15
    ph = fc.transformations[1]
    assert ph.gate_info.kind == G.GateKind.P and math.isclose(ph.gate_info.arg, 0.75)

    qs = QuantumCircuit(n)
    qs.h(1); qs.cp(2, 1, 0.3)        # H then CP on adjacent qubits -> fused pair (an uncontrolled
                                     # H then P is no special pair: can_fuse_special_pair, :181-219)
    qs.cp(0, 1, 0.1); qs.cp(0, 3, 0.2)   # CP.CP with a shared control -> fused pair
    qs.h(3)
    qs.rx(2, 0.4); qs.rx(2, 0.5)     # same target, not P -> UnitaryTransformation(m = 1)
    fs = fusion.fuse_circuit_specialized_pairs(qs)
    kinds = [type(t).__name__ for t in fs.transformations]
    assert kinds == ["FusedPairTransformation", "FusedPairTransformation", "GateTransformation",
                     "UnitaryTransformation"]
    assert fs.transformations[3].name == "fused" and fs.transformations[3].m == 1


def test_matrix_helpers():
    g = G.RY(0.7) @ G.RZ(0.2)
    assert np.array_equal(fusion.gate_to_matrix4x4(g, 2, 2), np.kron(np.eye(2), g))
    assert np.array_equal(fusion.gate_to_matrix4x4(g, 2, 3), np.kron(g, np.eye(2)))
    # controlled X with the control on the lower qubit: |b1 b0> -> |b1^b0 b0>
    cx = fusion.controlled_gate_to_matrix4x4(G.X, 0, 0, 1)
    perm = np.zeros((4, 4))
    for b0 in range(2):
        for b1 in range(2):
            perm[b0 + 2 * (b1 ^ b0), b0 + 2 * b1] = 1
    assert np.array_equal(cx, perm)


@pytest.mark.parametrize("which", ["mix", "qft", "value", "layered"])
@pytest.mark.parametrize("pass_", [fusion.fuse_circuit, fusion.fuse_circuit_specialized_pairs])
def test_fused_schedule_replays_to_the_oracle(which, pass_):
    n = 9
    fc = pass_(circuits(n)[which])
    ops, ud = lower_circuit(fc)
    psi = W.normal_state(n, 4).to_complex()
    plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=6))
    text = plan.dump()
    assert "unitary" not in text.split("\n", 1)[1]   # small blocks became gate records
    ref = oracle_run(n, fc, psi)
    assert np.abs(run_plan(text, psi, ud) - ref).max() <= 1e-12


def test_non_product_blocks_stay_unitary_steps():
    # a 4x4 that is no Kronecker product (a CX inside) must not be factored
    n = 4
    qc = QuantumCircuit(n)
    qc.transformations.append(UnitaryTransformation(fusion.controlled_gate_to_matrix4x4(G.X, 1, 1, 2), 1, 2))
    qc.transformations.append(UnitaryTransformation(np.kron(G.RY(0.3), G.H), 2, 2))
    ops, ud = lower_circuit(qc)
    text = Plan(n, np.float64, ops, ud, _lib.opts()).dump()
    assert text.count("unitary") == 1
    psi = W.normal_state(n, 1).to_complex()
    assert np.abs(run_plan(text, psi, ud) - oracle_run(n, qc, psi)).max() <= 1e-13
