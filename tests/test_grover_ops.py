"""Phase flips and diffusers (no GPU): the oracle's PHASE_FLIP / DIFFUSE
against dense matrices and Grover closed forms, the reference's classical
replacement rules, and the planner's Grover peephole replayed against the
gate-by-gate oracle."""
import math

import numpy as np
import pytest

from code_mojo_amd import _lib, algos, lower_circuit, replace_quantum_with_classical
from code_mojo_amd import ClassicalReplacementKind as CRK
from code_mojo_amd import workloads as W
from code_mojo_amd.circuit import ClassicalTransformation, QuantumCircuit
from code_mojo_amd.lowering import OP_DIFFUSE, OP_PHASE_FLIP, LoweringError, OpList
from code_mojo_amd.state import Plan
from oracle import oracle as O

from .plan_sim import run_plan

H2 = np.array([[1, 1], [1, -1]]) / math.sqrt(2)


def run(n, ops, ud, psi):
    re, im = psi.real.copy(), psi.imag.copy()
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def dense_on(n, mask, op_small):
    """Embed the 2^t x 2^t operator `op_small` on the bits of `mask`."""
    bits = [b for b in range(n) if (mask >> b) & 1]
    N = 1 << n
    M = np.zeros((N, N), dtype=complex)
    for col in range(N):
        e = sum(((col >> b) & 1) << i for i, b in enumerate(bits))
        base = col & ~mask
        for r in range(1 << len(bits)):
            row = base | sum(((r >> i) & 1) << b for i, b in enumerate(bits))
            M[row, col] += op_small[r, e]
    return M


@pytest.mark.parametrize("n,mask", [(3, 0b111), (4, 0b0110), (5, 0b10101), (5, 0b11111), (6, 0b110001)])
def test_diffuse_gate_variant_is_H_X_MCP_X_H(n, mask):
    t = bin(mask).count("1")
    Ht = np.array([[1.0]])
    for _ in range(t):
        Ht = np.kron(H2, Ht)
    flip0 = np.eye(1 << t)
    flip0[0, 0] = -1
    D = dense_on(n, mask, Ht @ flip0 @ Ht)   # H^t (I - 2|0><0|) H^t on the masked bits
    out = OpList(n)
    out.diffuse(mask, 0)
    ops, ud = out.finish()
    psi = rand_state(n, mask)
    assert np.abs(run(n, ops, ud, psi) - D @ psi).max() < 1e-13


@pytest.mark.parametrize("n,mask,value", [(3, 0b111, 5), (4, 0b1010, 0b1000), (5, 0b00110, 0), (6, 0b111111, 0)])
def test_phase_flip_is_diagonal_sign(n, mask, value):
    out = OpList(n)
    out.phase_flip(mask, value)
    ops, ud = out.finish()
    psi = rand_state(n, value + 11)
    idx = np.arange(1 << n)
    want = np.where((idx & mask) == value, -psi, psi)
    assert np.array_equal(run(n, ops, ud, psi), want)   # exact sign flips


@pytest.mark.parametrize("n", [3, 5, 7])
def test_classical_grover_closed_form(n):
    """Grover with the reference's classical diffuser from |s>: the marked
    probability after k iterations is sin^2((2k+1) theta), theta = asin(2^-n/2)
    (grover_classical.mojo:146-168); the classical diffuser equals minus the
    gate diffuser on real states, so the two runs differ by (-1)^k."""
    marked = (5 * n + 3) % (1 << n)
    iters = int(math.floor(math.pi / 4 * math.sqrt(2 ** n)))
    gate_qc = algos.grover_search_circuit(n, marked, iters)
    cls_qc = algos.grover_search_circuit(n, marked, iters)
    k = replace_quantum_with_classical(cls_qc, [CRK.GROVER_DIFFUSER])
    assert k == iters
    g_ops, g_ud = lower_circuit(gate_qc)
    c_ops, c_ud = lower_circuit(cls_qc)
    assert (c_ops["kind"] == OP_DIFFUSE).sum() == iters and (c_ops["arg_i"][c_ops["kind"] == OP_DIFFUSE] == 1).all()
    z = np.zeros(1 << n, complex)
    z[0] = 1
    g = run(n, g_ops, g_ud, z)
    c = run(n, c_ops, c_ud, z)
    assert np.abs(c - (-1) ** iters * g).max() < 1e-12
    theta = math.asin(2 ** (-n / 2))
    assert abs(abs(c[marked]) ** 2 - math.sin((2 * iters + 1) * theta) ** 2) < 1e-12


def test_inversion_0_rule_and_lowering():
    n = 4
    qc = QuantumCircuit(n)
    for q in range(n):
        qc.x(q)
    qc.mcp(list(range(n - 1)), n - 1, math.pi)
    for q in range(n):
        qc.x(q)
    gate_ops, gud = lower_circuit(qc)
    assert replace_quantum_with_classical(qc, [CRK.GROVER_INVERSION_0]) == 1
    assert isinstance(qc.transformations[0], ClassicalTransformation)
    assert qc.transformations[0].name == "GROVER_INVERSION_0"
    ops, ud = lower_circuit(qc)
    assert ops.size == 1 and ops[0]["kind"] == OP_PHASE_FLIP and ops[0]["aux"] == 0b1111 and ops[0]["controls"] == 0
    psi = rand_state(n, 4)
    assert np.abs(run(n, ops, ud, psi) - run(n, gate_ops, gud, psi)).max() < 1e-15


def test_rules_default_and_unknown_kind():
    qc = QuantumCircuit(3)
    qc.swap(0, 2)
    qc.qubit_reversal()
    qc.h(0)
    assert replace_quantum_with_classical(qc) == 2
    assert [getattr(t, "name", None) for t in qc.transformations[:2]] == ["SWAP", "BITREV"]
    with pytest.raises(ValueError, match="Unknown classical replacement kind"):
        replace_quantum_with_classical(qc, [9])


def test_sign_flip_oracle_and_target_checks():
    n = 3
    qc = QuantumCircuit(n)
    qc.add_classical("sign_flip_oracle", [5, 2, 5, 6])
    ops, ud = lower_circuit(qc)
    assert (ops["kind"] == OP_PHASE_FLIP).all() and list(ops["controls"]) == [5, 2, 5, 6]
    psi = rand_state(n, 1)
    want = psi.copy()
    for i in (5, 2, 5, 6):
        want[i] = -want[i]
    assert np.array_equal(run(n, ops, ud, psi), want)
    bad = QuantumCircuit(n)
    bad.add_classical("GROVER_DIFFUSER", [0, 0])
    with pytest.raises(LoweringError, match="Duplicate target"):
        lower_circuit(bad)


@pytest.mark.parametrize("n,tile_bits", [(8, 4), (9, 5), (12, 6)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_grover_peephole_schedule(n, tile_bits, dtype):
    """Above one tile, Grover's oracle and diffuser become one flip and one
    diffuser step each; the replayed schedule equals the gate-by-gate oracle."""
    qc, marked = W.grover_search(n, 17 + n)
    ops, ud = lower_circuit(qc)
    iters = int(math.floor(math.pi / 4 * math.sqrt(2 ** n)))
    plan = Plan(n, dtype, ops, ud, _lib.opts(tile_bits=tile_bits))
    text = plan.dump()
    assert text.count("\ndiffuse ") == iters
    # the initial H layer takes a few tile passes (the first oracle may join
    # them); every later oracle is absorbed into the preceding diffuser's pass
    # (chain flag 8) or stays a flip step
    absorbed = sum(1 for ln in text.split("\n") if ln.startswith("diffuse ") and int(ln.split("chain=")[1].split()[0]) & 8)
    assert iters - 1 <= text.count("\nflip ") + absorbed <= iters and text.count("\ntile ") <= n
    assert absorbed >= iters - 2
    z = np.zeros(1 << n, complex)
    z[0] = 1
    ref = run(n, ops, ud, z)
    got = run_plan(text, z, ud)
    assert np.abs(got - ref).max() < 1e-12
    assert abs(got[marked]) ** 2 > 0.9


def test_one_tile_grover_keeps_gates_but_folds_oracles():
    n = 9
    qc, _ = W.grover_search(n, 3)
    ops, ud = lower_circuit(qc)
    plan = Plan(n, np.float64, ops, ud, _lib.opts())
    text = plan.dump()
    assert "diffuse" not in text and plan.stats()["passes"] == 1
    z = np.zeros(1 << n, complex)
    z[0] = 1
    assert np.abs(run_plan(text, z, ud) - run(n, ops, ud, z)).max() < 1e-12


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_flip_and_diffuse_ops_inside_random_schedules(seed):
    """Random gate lists with phase flips and diffusers mixed in: flips are
    placed inside tile passes with patterned controls, diffusers split passes."""
    from .helpers import random_ops
    n = 9
    rng = np.random.default_rng(seed)
    ops, ud = random_ops(rng, n, 60, with_blocks=False, with_perms=False)
    extra = OpList(n)
    for _ in range(8):
        mask = int(rng.integers(1, 1 << n))
        extra.phase_flip(mask, int(rng.integers(0, 1 << n)) & mask)
    m2 = int(rng.integers(1, 1 << n)) & 0b1111111
    extra.diffuse(m2 or 1, int(rng.integers(0, 2)))
    eops, _ = extra.finish()
    allops = np.concatenate([ops[:20], eops[:4], ops[20:40], eops[4:], ops[40:]])
    psi = W.normal_state(n, seed).to_complex()
    ref = run(n, allops, ud, psi)
    for tb in (5, 7):
        plan = Plan(n, np.float64, allops, ud, _lib.opts(tile_bits=tb))
        assert np.abs(run_plan(plan.dump(), psi, ud) - ref).max() < 1e-12


@pytest.mark.parametrize("n,targets", [(16, list(range(14))), (20, [1, 3, 4, 7, 8, 9, 11, 12, 13, 14, 15, 16, 17, 18, 19]),
                                       (6, [0, 2, 3, 5])])
def test_partial_diffuser_of_any_width_lowers_and_plans(n, targets):
    """GROVER_DIFFUSER takes any target list (algos/grover_classical.mojo:284-326):
    diffusers over more than 12 of the qubits lower to one DIFFUSE op and the
    planner keeps it as one diffuser step; at 6 qubits the oracle equals the
    dense reflection of every block about its mean, per variant 1."""
    qc = QuantumCircuit(n)
    qc.add_classical("GROVER_DIFFUSER", targets)
    ops, ud = lower_circuit(qc)
    mask = sum(1 << t for t in targets)
    assert len(ops) == 1 and ops[0]["kind"] == OP_DIFFUSE and int(ops[0]["aux"]) == mask and ops[0]["arg_i"] == 1
    plan = Plan(n, np.float64, ops, ud, _lib.opts())
    assert f"diffuse mask={mask} variant=1" in plan.dump()
    if n <= 8:
        t = len(targets)
        J = np.full((1 << t, 1 << t), 2.0 / (1 << t))      # 2 |s><s| over the block
        D = dense_on(n, mask, J)
        psi = rand_state(n, 4)
        # variant 1: re <- 2 mean_re - re, im <- -2 mean_im - im
        want = (D @ psi.real - psi.real) + 1j * (-(D @ psi.imag) - psi.imag)
        assert np.abs(run(n, ops, ud, psi) - want).max() < 1e-14
