"""CPU tests of the host planner inside libbfly_hip.so (no device calls):
every compiled schedule, replayed by the numpy plan interpreter, must equal
the oracle on the same inputs, and planning must terminate with sane counts."""
import numpy as np
import pytest

from code_mojo_amd import _lib, algos, lower_circuit
from code_mojo_amd import gates as G
from code_mojo_amd import workloads as W
from code_mojo_amd.lowering import OpList
from code_mojo_amd.state import Plan
from oracle import oracle as O

from .helpers import random_ops
from .plan_sim import run_plan


def oracle_run(n, ops, ud, psi):
    re, im = psi.real.copy(), psi.imag.copy()
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


CASES = [(3, 0, 0), (4, 3, 0), (5, 4, 0), (5, 4, 1), (6, 0, 3), (7, 5, 2), (9, 6, 0), (10, 8, 4),
         (12, 8, 0), (12, 12, 0), (13, 10, 6)]


@pytest.mark.parametrize("n,tile_bits,low_bits", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fused_schedule_matches_oracle(n, tile_bits, low_bits, dtype):
    rng = np.random.default_rng(7 * n + tile_bits)
    ops, ud = random_ops(rng, n, 80)
    psi = W.normal_state(n, n).to_complex()
    plan = Plan(n, dtype, ops, ud, _lib.opts(tile_bits=tile_bits, low_bits=low_bits))
    got = run_plan(plan.dump(), psi, ud)
    ref = oracle_run(n, ops, ud, psi)
    assert np.abs(got - ref).max() < 1e-12
    st = plan.stats()
    assert 1 <= st["passes"] <= len(ops)


@pytest.mark.parametrize("max_gates", [1, 3, 17])
def test_gate_cap_respected(max_gates):
    n = 10
    rng = np.random.default_rng(max_gates)
    ops, ud = random_ops(rng, n, 60, with_blocks=False, with_perms=False)
    plan = Plan(n, np.float64, ops, ud, _lib.opts(max_gates_per_pass=max_gates))
    text = plan.dump()
    # a factored one-qubit gate is up to three instructions (phase, rotation,
    # phase), plus the pass's scale instructions
    for block in text.split("tile ")[1:]:
        assert block.count("instr ") <= 3 * max_gates + 2
    # adjacent gates with the same target and controls merge into one (runs
    # of at most 64), so the cap counts runs
    runs = sum(1 for i, o in enumerate(ops)
               if not (i and o["kind"] == 0 and ops[i - 1]["kind"] == 0 and o["target"] == ops[i - 1]["target"]
                       and o["controls"] == ops[i - 1]["controls"]))
    assert plan.stats()["passes"] >= -(-runs // max_gates)
    psi = W.normal_state(n, 3).to_complex()
    assert np.abs(run_plan(text, psi, ud) - oracle_run(n, ops, ud, psi)).max() < 1e-12


def test_unfused_schedule_matches_oracle():
    n = 8
    rng = np.random.default_rng(5)
    ops, ud = random_ops(rng, n, 60)
    plan = Plan(n, np.float64, ops, ud, _lib.opts(mode=_lib.MODE_UNFUSED))
    psi = W.normal_state(n, 4).to_complex()
    assert np.abs(run_plan(plan.dump(), psi, ud) - oracle_run(n, ops, ud, psi)).max() < 1e-12


def test_config_circuits_pass_counts():
    """QFT: H gates are the only non-diagonal ops, so a 12-qubit tile covers a
    20-qubit QFT in ceil((20-12)/(12-4)) + 1 = 2 passes plus the bit reversal;
    layered circuits need about three tile passes per layer at 28 qubits."""
    ops, ud = lower_circuit(W.qft_circuit(20))
    st = Plan(20, np.float64, ops, ud, _lib.opts()).stats()
    assert st["passes"] == 3, st
    ops, ud = lower_circuit(W.random_layered(28, 4, seed=1))
    st = Plan(28, np.float64, ops, ud, _lib.opts()).stats()
    assert st["passes"] <= 4 * 3 + 1, st
    ops, ud = lower_circuit(algos.encode_value_circuit(30, 123456789, exact_angles=True, swap=True))
    st = Plan(30, np.float64, ops, ud, _lib.opts()).stats()
    assert st["passes"] <= 5, st


def test_layered_schedule_matches_oracle():
    n = 12
    qc = W.random_layered(n, 6, seed=9)
    ops, ud = lower_circuit(qc)
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    for tb in (0, 6, 9):
        plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb))
        assert np.abs(run_plan(plan.dump(), psi, ud) - oracle_run(n, ops, ud, psi)).max() < 1e-12


def test_grover_and_ae_schedules():
    n = 7
    qc, marked = W.grover_search(n, 3)
    ops, ud = lower_circuit(qc)
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=5))
    got = run_plan(plan.dump(), psi, ud)
    assert np.abs(got - oracle_run(n, ops, ud, psi)).max() < 1e-12
    qc, a = W.amplitude_estimation(4)
    ops, ud = lower_circuit(qc)
    n = qc.num_qubits
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=6))
    assert np.abs(run_plan(plan.dump(), psi, ud) - oracle_run(n, ops, ud, psi)).max() < 1e-12


def test_invalid_ops_rejected():
    n = 4
    ops, ud = lower_circuit(W.qft_circuit(n))
    bad = ops.copy()
    bad["target"][0] = 9
    with pytest.raises(_lib.BflyError, match="Target out of bounds"):
        Plan(n, np.float64, bad, ud, _lib.opts())
    bad = ops.copy()
    bad["controls"][1] = 1 << int(bad["target"][1])
    with pytest.raises(_lib.BflyError, match="overlaps"):
        Plan(n, np.float64, bad, ud, _lib.opts())
    bad = ops.copy()
    bad["kind"][0] = 42
    with pytest.raises(_lib.BflyError):
        Plan(n, np.float64, bad, ud, _lib.opts())
    bad = ops.copy()
    bad["mat"][0, 0] = np.nan
    with pytest.raises(_lib.BflyError, match="non-finite"):
        Plan(n, np.float64, bad, ud, _lib.opts())
    with pytest.raises(_lib.BflyError, match="execution mode"):
        Plan(n, np.float64, ops, ud, _lib.opts(mode=7))


def _u3_layer(ops, rng, n, qubits=None):
    for q in (range(n) if qubits is None else qubits):
        g = G.Custom_Gate(G.U3(*(float(x) for x in rng.uniform(0.1, 3.0, size=3))))
        ops.gate(q, [], g.kind, g.gate, 0.0)


def test_trailing_phases_ride_on_later_rotations():
    """A one-qubit unitary factors into phase, real rotation, phase; the last
    phase of each qubit in a tile pass waits for the qubit's next rotation in
    a later pass (its pre-phase) instead of becoming an instruction of its
    own.  Counted on a 14-qubit layered circuit with 2^6 tiles (many passes per
    layer), and replayed against the oracle."""
    n = 14
    ops, ud = lower_circuit(W.random_layered(n, 8, seed=4))
    plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=6))
    text = plan.dump()
    rot = text.count("instr kind=3 ") + text.count("instr kind=6 ")
    ph = text.count("instr kind=4 ") + text.count("instr kind=2 ")
    assert plan.stats()["passes"] > 8
    assert ph * 4 <= rot, (ph, rot)
    psi = W.normal_state(n, 2).to_complex()
    assert np.abs(run_plan(text, psi, ud) - oracle_run(n, ops, ud, psi)).max() < 1e-12


@pytest.mark.parametrize("barrier", ["bitrev", "diffuse", "measure", "end"])
def test_waiting_phases_are_applied_before_a_barrier(barrier):
    """Phases that wait for a later pass are applied before any op that does
    not commute with them (a bit reversal, a diffuser, a measurement) and by
    the end of the circuit, even when the qubit has more rotations after it."""
    n = 10
    rng = np.random.default_rng(11)
    ops = OpList(n)
    for _ in range(3):
        _u3_layer(ops, rng, n)
        for a in range(0, n - 1, 2):
            ops.gate(a + 1, [a], G.GateKind.names.index("Z"), G.GateInfo.of(G.GateKind.names.index("Z"), None).gate, 0.0)
    if barrier == "bitrev":
        ops.bitrev((1 << n) - 1)
    elif barrier == "diffuse":
        ops.diffuse((1 << n) - 1, 0)
    elif barrier == "measure":
        ops.measure([0, 5], [None, None], seed=12345)
    if barrier != "end":
        _u3_layer(ops, rng, n, qubits=[1, 4, 7])
    ops, ud = ops.finish()
    psi = W.normal_state(n, 6).to_complex()
    ref = oracle_run(n, ops, ud, psi)
    for tb in (4, 6):
        plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb))
        got = run_plan(plan.dump(), psi, ud)
        assert np.abs(got - ref).max() < 1e-12, (barrier, tb)


def test_long_single_tile_circuits_get_narrower_tiles():
    """A state that fits one default tile would run each pass in one
    workgroup: a long circuit (>= 48 gates per qubit, here a 12-qubit Grover
    search) compiles to tiles two bits narrower than the state, so its
    iterations become phase flips and diffusers; a short one (a 12-qubit,
    4-layer circuit) keeps the single tile; an explicit tile size is kept."""
    from code_mojo_amd import serve
    job = serve.grover_job(12, 3, np.float32)
    text = Plan(12, np.float32, job.ops, job.udata, _lib.opts()).dump()
    assert "tile k=10 " in text and "diffuse" in text
    ops, ud = lower_circuit(W.random_layered(12, 4, seed=5))
    assert "tile k=12 " in Plan(12, np.float32, ops, ud, _lib.opts()).dump()
    assert "tile k=12 " in Plan(12, np.float32, job.ops, job.udata, _lib.opts(tile_bits=12)).dump()
    psi = np.zeros(1 << 12, dtype=complex)
    psi[0] = 1
    got = run_plan(text, psi, job.udata)
    assert np.abs(got - oracle_run(12, job.ops, job.udata, psi)).max() < 1e-12


def test_tile_search_cuts_passes_and_matches_first_fit():
    """The planner's tile-qubit search (default) and first fit (BFLY_FLAG_FIRST_FIT)
    give the same amplitudes; on layered circuits the search needs far fewer
    passes (the 30-qubit headline circuit: 55 passes first fit, 30 searched)."""
    ops, ud = lower_circuit(W.random_layered(30, 40, 2026))
    ff = Plan(30, np.float64, ops, ud, _lib.opts(flags=_lib.FLAG_FIRST_FIT)).stats()["passes"]
    se = Plan(30, np.float64, ops, ud, _lib.opts()).stats()["passes"]
    assert ff == 55 and se <= 32, (ff, se)
    n = 14
    qc = W.random_layered(n, 12, seed=4)
    ops, ud = lower_circuit(qc)
    psi = W.normal_state(n, 2).to_complex()
    ref = oracle_run(n, ops, ud, psi)
    for tb in (6, 8, 10):
        for fl in (0, _lib.FLAG_FIRST_FIT):
            plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb, flags=fl))
            assert np.abs(run_plan(plan.dump(), psi, ud) - ref).max() < 1e-12
        assert (Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb)).stats()["passes"]
                <= Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb, flags=_lib.FLAG_FIRST_FIT)).stats()["passes"])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tile_search_random_circuits_match_oracle(seed):
    """Random gate lists (controls, diagonal and dense ops) on 11 qubits with
    2^5-2^7 tiles: the searched schedules replay to the oracle's amplitudes."""
    n = 11
    rng = np.random.default_rng(100 + seed)
    ops, ud = random_ops(rng, n, 150)
    psi = W.normal_state(n, seed).to_complex()
    ref = oracle_run(n, ops, ud, psi)
    for tb in (5, 6, 7):
        plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb))
        assert np.abs(run_plan(plan.dump(), psi, ud) - ref).max() < 1e-12


def test_runs_on_one_target_merge_into_one_gate():
    """Adjacent gates with the same target and controls multiply into one 2x2
    gate (runs of at most 64): amplitude estimation's 1569 controlled gates on
    qubit 0 become a few dozen tile instructions, and long runs of every gate
    kind (200 gates, controlled and not) replay to the oracle's amplitudes."""
    from code_mojo_amd import serve
    job = serve.ae_job(5, np.float64)
    text = Plan(job.n, np.float64, job.ops, job.udata, _lib.opts()).dump()
    assert text.count("instr ") < 120, text.count("instr ")
    n = 8
    rng = np.random.default_rng(21)
    ops = OpList(n)
    kinds = ["H", "P", "X", "Y", "Z", "RX", "RY", "RZ"]
    for target, ctrls in ((0, []), (3, [5]), (7, [1, 2]), (2, [])):
        for _ in range(200):
            k = G.GateKind.names.index(kinds[int(rng.integers(len(kinds)))])
            info = G.GateInfo.of(k, float(rng.uniform(-3, 3)) if k in (1, 5, 6, 7) else None)
            ops.gate(target, ctrls, info.kind, info.gate, info.arg or 0.0)
    ops, ud = ops.finish()
    psi = W.normal_state(n, 4).to_complex()
    ref = oracle_run(n, ops, ud, psi)
    for tb in (0, 5):
        plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb))
        assert plan.dump().count("instr ") < 80
        assert np.abs(run_plan(plan.dump(), psi, ud) - ref).max() < 1e-12
    # the reference's per-gate arithmetic (generic pairs) keeps every gate
    plan = Plan(n, np.float64, ops, ud, _lib.opts(flags=_lib.FLAG_GENERIC_PAIRS))
    assert plan.dump().count("instr ") >= 700   # (consecutive sign flips still XOR-merge)
    assert np.abs(run_plan(plan.dump(), psi, ud) - ref).max() < 1e-12


def test_beam_schedules_match_oracle_and_need_no_more_passes(monkeypatch):
    """The beam over whole schedules (BFLY_PLAN_BEAM=1) replays to the oracle's
    amplitudes and never needs more passes than the greedy search (the
    30-qubit headline circuit: 28 against 30)."""
    ops, ud = lower_circuit(W.random_layered(30, 40, 2026))
    greedy = Plan(30, np.float64, ops, ud, _lib.opts()).stats()["passes"]
    monkeypatch.setenv("BFLY_PLAN_BEAM", "1")
    beam = Plan(30, np.float64, ops, ud, _lib.opts()).stats()["passes"]
    assert beam <= 28 and beam <= greedy, (beam, greedy)
    n = 14
    rng = np.random.default_rng(8)
    for qc in (W.random_layered(n, 16, seed=6), None):
        if qc is None:
            ops, ud = random_ops(rng, n, 200)
        else:
            ops, ud = lower_circuit(qc)
        psi = W.normal_state(n, 7).to_complex()
        ref = oracle_run(n, ops, ud, psi)
        for tb in (6, 8, 10):
            monkeypatch.setenv("BFLY_PLAN_BEAM", "1")
            plan = Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb))
            assert np.abs(run_plan(plan.dump(), psi, ud) - ref).max() < 1e-12
            monkeypatch.setenv("BFLY_PLAN_BEAM", "0")
            assert plan.stats()["passes"] <= Plan(n, np.float64, ops, ud, _lib.opts(tile_bits=tb)).stats()["passes"]
