"""Measurement (no GPU): the oracle's apply_measure restatement against what
the mathematics fixes (uniform states collapse to the outcome the MINSTD draw
selects, post-selection is the normalised projection, the reference's
seed-0 and zero-probability errors), an independent numpy restatement, the
lowering of MeasurementTransformation, and plan replays with measurements
in the middle of circuits."""
import math

import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit, measure
from code_mojo_amd import workloads as W
from code_mojo_amd.circuit import MeasurementTransformation, QuantumCircuit
from code_mojo_amd.context import ExecContext
from code_mojo_amd.lowering import OP_MEASURE, LoweringError, OpList
from code_mojo_amd.state import Plan
from oracle import oracle as O

from .plan_sim import apply_measure, run_plan

M = 2147483647


def oracle_measure(psi, n, targets, values, seed, dtype=np.float64):
    out = OpList(n)
    out.measure(targets, [None if v < 0 else bool(v) for v in values], seed)
    ops, ud = out.finish()
    re, im = psi.real.astype(dtype).copy(), psi.imag.astype(dtype).copy()
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("seed", [1, 7, 12345, 2 ** 40 + 3, 987654321])
def test_uniform_state_collapses_to_the_drawn_outcome(seed):
    """|+>^3: every outcome has p = 1/8; the chosen one is the first o with
    r <= (o + 1) / 8, r = ((seed mod M) * 48271 mod M) / M."""
    n = 3
    psi = np.full(8, 1 / math.sqrt(8), dtype=complex)
    rng = (seed % M) * 48271 % M
    r = rng / M
    want = next(o for o in range(8) if r <= (o + 1) / 8)
    got = oracle_measure(psi, n, [0, 1, 2], [-1, -1, -1], seed)
    expect = np.zeros(8, complex)
    expect[want] = 1
    assert np.abs(got - expect).max() < 1e-15


def test_outcome_bit_order_follows_the_target_list():
    """Free targets [2, 0]: outcome bit 0 is qubit 2.  On |100> (index 4) the
    only outcome with probability is 0b01."""
    n = 3
    psi = np.zeros(8, complex)
    psi[4] = 1
    out = OpList(n)
    out.measure([2, 0], [None, None], 5)
    ops, ud = out.finish()
    re, im = psi.real.copy(), psi.imag.copy()
    O.execute(n, re, im, ops, ud)
    assert re[4] == 1.0 and np.count_nonzero(re) == 1


@pytest.mark.parametrize("n,target,value", [(4, 1, True), (5, 3, False), (6, 0, True)])
def test_forced_value_is_the_normalised_projection(n, target, value):
    psi = rand_state(n, n + target)
    idx = np.arange(1 << n)
    keep = ((idx >> target) & 1) == int(value)
    proj = np.where(keep, psi, 0)
    want = proj / np.linalg.norm(proj)
    got = oracle_measure(psi, n, [target], [int(value)], 99)
    assert np.abs(got - want).max() < 1e-14


@pytest.mark.parametrize("seed", [3, 11, 2024, 77777])
def test_oracle_matches_numpy_restatement(seed):
    n = 7
    psi = rand_state(n, seed)
    targets, values = [5, 1, 3, 6], [-1, 1, -1, -1]
    got = oracle_measure(psi, n, targets, values, seed)
    want = apply_measure(psi, n, targets, values, seed)
    assert np.abs(got - want).max() < 1e-13
    assert abs(np.linalg.norm(got) - 1) < 1e-13


def test_complex64_oracle_matches_within_float_rounding():
    n = 8
    psi = rand_state(n, 5)
    got = oracle_measure(psi, n, [0, 4], [-1, -1], 31, dtype=np.float32)
    want = apply_measure(psi.astype(np.complex64).astype(complex), n, [0, 4], [-1, -1], 31)
    assert np.abs(got - want).max() < 2e-6


def test_reference_errors():
    n = 2
    zero = np.zeros(4, complex)
    zero[0] = 1
    with pytest.raises(O.OracleError, match="Measurement probability is zero"):
        oracle_measure(zero, n, [0], [1], 5)           # forced |1> on |0>
    one = np.zeros(4, complex)
    one[1] = 1
    with pytest.raises(O.OracleError, match="outcome probability is zero"):
        oracle_measure(one, n, [0], [-1], 0)           # seed 0: r = 0 picks outcome 0 (p = 0)


def test_planted_mistakes_fail():
    """A reversed outcome bit order, or the older MINSTD multiplier 16807,
    changes the collapsed state for most seeds."""
    n = 6
    psi = rand_state(n, 8)
    targets, values = [0, 3, 5], [-1, -1, -1]
    inv48271 = pow(48271, -1, M)
    order = mult = 0
    # seeds spread over [1, M): a small seed gives r ~ seed * 48271 / M, i.e.
    # almost always outcome 0 (a property of the reference's single step)
    for seed in np.random.default_rng(0).integers(1, M, 40).tolist():
        good = oracle_measure(psi, n, targets, values, seed)
        order += np.abs(good - apply_measure(psi, n, targets[::-1], values, seed)).max() > 1e-3
        bad_seed = seed * 16807 * inv48271 % M       # r drawn with multiplier 16807
        mult += np.abs(good - apply_measure(psi, n, targets, values, bad_seed)).max() > 1e-3
    assert order >= 10 and mult >= 10, (order, mult)


def test_lowering_and_config_seed():
    qc = QuantumCircuit(4)
    qc.h(0)
    measure(qc, [0, 2], [None, True], seed=17)
    measure(qc, [1])
    ops, ud = lower_circuit(qc, measure_seed=5)
    ms = ops[ops["kind"] == OP_MEASURE]
    assert ms.size == 2
    assert ms[0]["gate"] == 1 and ms[0]["controls"] == 17 and ms[0]["arg_i"] == 2
    assert list(ud[2 * ms[0]["aux"]: 2 * ms[0]["aux"] + 4]) == [0, -1, 2, 1]
    assert ms[1]["gate"] == 1 and ms[1]["controls"] == 5                   # configured seed
    ops2, _ = lower_circuit(qc)
    assert ops2[ops2["kind"] == OP_MEASURE][1]["gate"] == 0               # clock at run time
    assert ExecContext.from_config({"measure_seed": "9"}).measure_seed == 9
    bad = QuantumCircuit(3)
    bad.transformations.append(MeasurementTransformation([0, 1], 3, [True]))
    with pytest.raises(LoweringError, match="values size mismatch"):
        lower_circuit(bad)
    oob = QuantumCircuit(3)
    measure(oob, [3])
    with pytest.raises(LoweringError, match="Target out of bounds"):
        lower_circuit(oob)


@pytest.mark.parametrize("mode", [_lib.MODE_FUSED, _lib.MODE_UNFUSED])
def test_plan_with_mid_circuit_measurement_replays_to_the_oracle(mode):
    n = 10
    qc = W.random_layered(n, 3, 11)
    measure(qc, [2, 7, 9], [None, None, False], seed=123)
    for q in range(n):
        qc.h(q)
    measure(qc, [0, 1, 2, 3], seed=456)
    ops, ud = lower_circuit(qc)
    psi = W.normal_state(n, 2).to_complex()
    re, im = psi.real.copy(), psi.imag.copy()
    O.execute(n, re, im, ops, ud)
    plan = Plan(n, np.float64, ops, ud, _lib.opts(mode=mode))
    text = plan.dump()
    assert text.count("\nmeasure ") == 2
    got = run_plan(text, psi, ud)
    assert np.abs(got - (re + 1j * im)).max() < 1e-12
