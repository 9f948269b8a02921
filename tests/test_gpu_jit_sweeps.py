"""GPU parity of the generated tile-pass kernels' sweep forms, against the
oracle at the drop-in's bars (complex128 max |amplitude error| <= 1e-12;
complex64 against the complex128 oracle: L2 <= 2e-5 and fidelity >= 1 - 1e-6):

- 8-amplitude (3-slot) and 16-amplitude (4-slot) register groups at tile
  sizes 2^10..2^13, on the padded LDS layout;
- member factors: rotations' pre-phases, phases, sign flips and scale
  factors applied once per register member, and lane- or tile-conditional
  flips (a CZ with a lane or outer qubit) as sign XORs.  Circuits rich in
  X / Y (quarter-turn rotations, the cotangent form) and in CZ / controlled
  phases between random pairs exercise every folding case;
- direct stage-in / stage-out (the first sweep loads its members from HBM
  in its own layout, the last stores them back: scalar, 2- and 4-wide vector
  accesses), forced on every pass of every sweep form;
- the work-queue launch form (blocks take tiles from the executing state's
  counter and load the next tile during the current one), forced on every
  pass of every sweep form, repeated on a new state, and one plan on two
  states at once;
- split round trips (complex128 default: re and im through a one-component
  LDS tile in turn), in every form above, and bit-identical at 26 qubits to
  16-byte round trips;
- the measurement switches BFLY_JIT_HOIST=0 (every diagonal action in place),
  BFLY_JIT_GRID=persistent (blocks walk the tiles, with or without the next
  tile's loads during the sweeps: BFLY_JIT_PREFETCH) and BFLY_JIT_DIRECT
  (every pass staged through LDS, or direct on one or both sides) give the
  same amplitudes within the complex128 bar."""
import math
import os

import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit
from code_mojo_amd import gates as G
from code_mojo_amd import workloads as W
from code_mojo_amd.lowering import OpList
from code_mojo_amd.state import GpuState, Plan, QuantumState
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_C128 = 1e-12


def run(n, ops, ud, st, dtype=np.float64, **kw):
    g = GpuState(n, dtype)
    g.upload(QuantumState(n, dtype, st.re.astype(dtype), st.im.astype(dtype)))
    p = Plan(n, dtype, ops, ud, _lib.opts(flags=_lib.FLAG_JIT, **kw))
    info = p.jit_info()
    p.execute(g)
    out = g.download()
    p.free()
    g.free()
    return out.re.astype(np.float64) + 1j * out.im.astype(np.float64), info


def oracle(n, ops, ud, st):
    re, im = st.re.astype(np.float64).copy(), st.im.astype(np.float64).copy()
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


def flips_and_quarter_turns(n, count, seed):
    """X, Y, H, U3 and RX near quarter turns on random qubits, interleaved with
    CZ, CP and multi-controlled Z between random qubits (the sign and phase
    records that fold into member factors or lane-conditional XORs)."""
    rng = np.random.default_rng(seed)
    ops = OpList(n)
    for _ in range(count):
        r = rng.random()
        t = int(rng.integers(n))
        others = [q for q in range(n) if q != t]
        if r < 0.45:
            k = ["X", "Y", "H", "U3", "RX"][int(rng.integers(5))]
            if k == "U3":
                info = G.Custom_Gate(G.U3(float(rng.uniform(0, 6.3)), float(rng.uniform(0, 6.3)),
                                          float(rng.uniform(0, 6.3))))
            elif k == "RX":
                info = G.GateInfo.of(G.GateKind.RX, math.pi * (1 + 1e-9 * rng.normal()))
            else:
                info = G.GateInfo.of(G.GateKind.names.index(k))
            ops.gate(t, [], info.kind, info.gate, info.arg or 0.0)
        elif r < 0.75:
            c = int(rng.choice(others))
            info = G.GateInfo.of(G.GateKind.Z)
            ops.gate(t, [c], info.kind, info.gate, 0.0)
        elif r < 0.9:
            c = int(rng.choice(others))
            info = G.GateInfo.of(G.GateKind.P, float(rng.uniform(-math.pi, math.pi)))
            ops.gate(t, [c], info.kind, info.gate, info.arg)
        else:
            cs = [int(x) for x in rng.choice(others, size=min(2, len(others)), replace=False)]
            info = G.GateInfo.of(G.GateKind.Z)
            ops.gate(t, cs, info.kind, info.gate, 0.0)
    return ops.finish()


# (qubits, tile bits, threads): threads 2^(k-3) selects 3-slot sweeps, 2^(k-4) 4-slot
SWEEP_CASES = [(16, 10, 128), (16, 10, 64), (17, 11, 256), (17, 11, 128), (18, 12, 512), (18, 12, 256),
               (19, 13, 1024), (19, 13, 512)]


@pytest.mark.parametrize("n,tile_bits,threads", SWEEP_CASES)
def test_layered_sweep_forms_vs_oracle(gpu_lib, n, tile_bits, threads):
    qc = W.random_layered(n, 10, seed=n + threads)
    ops, ud = lower_circuit(qc)
    st = W.normal_state(n, n)
    out, info = run(n, ops, ud, st, tile_bits=tile_bits, threads=threads)
    assert info["kernels"] > 0
    err = np.abs(out - oracle(n, ops, ud, st)).max()
    assert err <= TOL_C128, (n, tile_bits, threads, err)


@pytest.mark.parametrize("n,tile_bits,threads", SWEEP_CASES)
def test_flip_heavy_circuits_vs_oracle(gpu_lib, n, tile_bits, threads):
    ops, ud = flips_and_quarter_turns(n, 400, seed=7 * n + threads)
    st = W.normal_state(n, 2 * n)
    out, info = run(n, ops, ud, st, tile_bits=tile_bits, threads=threads)
    assert info["kernels"] > 0
    err = np.abs(out - oracle(n, ops, ud, st)).max()
    assert err <= TOL_C128, (n, tile_bits, threads, err)


ALL_DIRECT = {"BFLY_JIT_DIRECT": "1", "BFLY_JIT_DIRECT_MIN_SWEEPS": "1"}
# the work-queue launch form (persistent blocks taking tiles from the state's
# counter, next tile loaded during the current one) on every pass
ALL_QUEUE = {"BFLY_JIT_QUEUE_MIN_TILES": "1", "BFLY_JIT_SPLIT": "0"}


@pytest.mark.parametrize("n,tile_bits,threads", SWEEP_CASES)
def test_direct_stage_in_and_out_vs_oracle(gpu_lib, monkeypatch, n, tile_bits, threads):
    for k, v in ALL_DIRECT.items():
        monkeypatch.setenv(k, v)
    ops, ud = flips_and_quarter_turns(n, 300, seed=3 * n + threads)
    st = W.normal_state(n, 4 * n)
    out, info = run(n, ops, ud, st, tile_bits=tile_bits, threads=threads)
    assert info["kernels"] > 0
    err = np.abs(out - oracle(n, ops, ud, st)).max()
    assert err <= TOL_C128, (n, tile_bits, threads, err)


ALL_SPLIT = {"BFLY_JIT_SPLIT_MIN_SWEEPS": "1"}
SPLIT_QUEUE = {"BFLY_JIT_SPLIT_MIN_SWEEPS": "1", "BFLY_JIT_GRID": "queue", "BFLY_JIT_PREFETCH": "1"}


@pytest.mark.parametrize("switches", [ALL_SPLIT, SPLIT_QUEUE], ids=["split", "split_queue_prefetch"])
@pytest.mark.parametrize("n,tile_bits,threads", SWEEP_CASES)
def test_split_round_trips_vs_oracle(gpu_lib, monkeypatch, n, tile_bits, threads, switches):
    """Every complex128 pass with re and im through a one-component LDS tile in
    turn; with the work queue and next-tile loads, staged passes issue those
    loads once the first sweep has read the staged im parts back."""
    for k, v in switches.items():
        monkeypatch.setenv(k, v)
    ops, ud = flips_and_quarter_turns(n, 300, seed=11 * n + threads)
    st = W.normal_state(n, 8 * n)
    out, info = run(n, ops, ud, st, tile_bits=tile_bits, threads=threads)
    assert info["kernels"] > 0
    err = np.abs(out - oracle(n, ops, ud, st)).max()
    assert err <= TOL_C128, (n, tile_bits, threads, err)


@pytest.mark.parametrize("n,tile_bits,threads", SWEEP_CASES)
def test_work_queue_form_vs_oracle(gpu_lib, monkeypatch, n, tile_bits, threads):
    for k, v in ALL_QUEUE.items():
        monkeypatch.setenv(k, v)
    ops, ud = flips_and_quarter_turns(n, 300, seed=5 * n + threads)
    st = W.normal_state(n, 6 * n)
    ref = oracle(n, ops, ud, st)
    out, info = run(n, ops, ud, st, tile_bits=tile_bits, threads=threads)
    assert info["kernels"] > 0
    assert np.abs(out - ref).max() <= TOL_C128, (n, tile_bits, threads)
    # the same plan again on a new state: the counters are cleared per execution
    out2, _ = run(n, ops, ud, st, tile_bits=tile_bits, threads=threads)
    assert np.array_equal(out, out2)


def test_one_work_queue_plan_on_two_states_at_once(gpu_lib, monkeypatch):
    """Calls on different states may overlap: each state holds its own tile
    counters, so one plan running on two states' streams at once gives both
    the oracle's amplitudes."""
    for k, v in ALL_QUEUE.items():
        monkeypatch.setenv(k, v)
    n = 19
    qc = W.random_layered(n, 12, seed=31)
    ops, ud = lower_circuit(qc)
    sts = [W.normal_state(n, 41), W.normal_state(n, 42)]
    gs = []
    for st in sts:
        g = GpuState(n, np.float64)
        g.upload(QuantumState(n, np.float64, st.re, st.im))
        gs.append(g)
    p = Plan(n, np.float64, ops, ud, _lib.opts(flags=_lib.FLAG_JIT))
    for _ in range(3):
        for g in gs:
            p.execute(g)      # asynchronous: the two states' streams overlap
    for g, st in zip(gs, sts):
        out = g.download()
        got = out.re + 1j * out.im
        re, im = st.re.copy(), st.im.copy()
        for _ in range(3):
            O.execute(n, re, im, ops, ud)
        assert np.abs(got - (re + 1j * im)).max() <= TOL_C128
        g.free()
    p.free()


@pytest.mark.parametrize("form", ["tile", "direct", "queue"])
@pytest.mark.parametrize("tile_bits,threads", [(13, 1024), (13, 512), (12, 256)])
def test_complex64_sweep_forms_vs_complex128_oracle(gpu_lib, monkeypatch, tile_bits, threads, form):
    for k, v in {"tile": {}, "direct": ALL_DIRECT, "queue": {**ALL_DIRECT, **ALL_QUEUE}}[form].items():
        monkeypatch.setenv(k, v)
    n = 19
    ops, ud = flips_and_quarter_turns(n, 300, seed=tile_bits + threads)
    st = W.normal_state(n, 5)
    out, _ = run(n, ops, ud, st, dtype=np.float32, tile_bits=tile_bits, threads=threads)
    ref = oracle(n, ops, ud, st)
    l2 = float(np.linalg.norm(out - ref))
    fid = float(abs(np.vdot(ref, out)) ** 2 / (np.vdot(ref, ref).real * np.vdot(out, out).real))
    assert l2 <= 2e-5 and fid >= 1 - 1e-6, (tile_bits, threads, l2, fid)


@pytest.mark.parametrize("switches", [
    {"BFLY_JIT_HOIST": "0"},
    {"BFLY_JIT_GRID": "persistent"}, {"BFLY_JIT_GRID": "persistent", "BFLY_JIT_PREFETCH": "1"},
    {"BFLY_JIT_GRID": "queue"}, {"BFLY_JIT_GRID": "queue", "BFLY_JIT_PREFETCH": "1"},
    {"BFLY_JIT_GRID": "queue", "BFLY_JIT_PREFETCH": "1", "BFLY_JIT_DIRECT": "1", "BFLY_JIT_DIRECT_MIN_SWEEPS": "1"},
    {"BFLY_JIT_DIRECT": "0"}, ALL_DIRECT,
    {"BFLY_JIT_DIRECT": "in", "BFLY_JIT_DIRECT_MIN_SWEEPS": "1"},
    {"BFLY_JIT_DIRECT": "out", "BFLY_JIT_DIRECT_MIN_SWEEPS": "1"},
    {"BFLY_JIT_SPLIT": "0"}, {"BFLY_JIT_SPLIT": "0", "BFLY_JIT_DIRECT": "0"},
    {"BFLY_JIT_SPLIT": "0", "BFLY_JIT_GRID": "queue"},
    {"BFLY_JIT_SPLIT": "0", "BFLY_JIT_GRID": "queue", "BFLY_JIT_PREFETCH": "0"},
    {"BFLY_JIT_SPLIT": "0", "BFLY_JIT_GRID": "persistent"},
])
def test_measurement_switches_give_the_same_amplitudes(gpu_lib, switches):
    n = 18
    qc = W.random_layered(n, 12, seed=11)
    ops, ud = lower_circuit(qc)
    st = W.normal_state(n, 3)
    ref = oracle(n, ops, ud, st)
    base, _ = run(n, ops, ud, st)
    old = {k: os.environ.get(k) for k in switches}
    os.environ.update(switches)
    try:
        off, info = run(n, ops, ud, st)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    assert info["compiled"] > 0   # the switched sources differ, so they compile anew
    assert np.abs(off - ref).max() <= TOL_C128
    assert np.abs(base - ref).max() <= TOL_C128


@pytest.mark.parametrize("dtype,other", [(np.float64, {"BFLY_JIT_SPLIT": "0"}),
                                         (np.float64, {"BFLY_JIT_SPLIT": "0", "BFLY_JIT_GRID": "tile"}),
                                         (np.float32, {"BFLY_JIT_GRID": "tile"})])
def test_default_launch_forms_are_bit_identical_at_26_qubits(gpu_lib, monkeypatch, dtype, other):
    """26 qubits (2^14 complex128 / 2^13 complex64 tiles per pass): complex128
    passes run split round trips one block per tile by default, complex64
    passes the work queue; every form applies the same arithmetic to each
    tile, so the amplitudes equal those of the other forms bit for bit."""
    n = 26
    qc = W.random_layered(n, 10, seed=77)
    ops, ud = lower_circuit(qc)
    st = W.normal_state(n, 9)
    q, info = run(n, ops, ud, st, dtype=dtype)
    assert info["kernels"] > 0
    for k, v in other.items():
        monkeypatch.setenv(k, v)
    t, info_t = run(n, ops, ud, st, dtype=dtype)
    assert info_t["compiled"] > 0      # a different form, compiled anew
    assert np.array_equal(q, t)
    nrm = float(np.vdot(q, q).real)
    assert abs(nrm - float(np.vdot(st.re + 1j * st.im, st.re + 1j * st.im).real)) <= (1e-9 if dtype == np.float64 else 1e-3)
