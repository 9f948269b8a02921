"""GPU: batched execution of many circuits (one launch per round per tile
size) must give every circuit exactly the result it gets alone; a captured
hipGraph must give the eager result on every replay."""
import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd.state import Batch, GpuState, Plan

from .helpers import random_ops

pytestmark = pytest.mark.gpu


def mixed_jobs(seed, count=12):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        n = int(rng.integers(4, 17))
        dtype = np.float64 if i % 2 == 0 else np.float32
        if i % 3 == 0:
            ops, ud = lower_circuit(W.random_layered(n, 4, seed + i))
        elif i % 3 == 1:
            ops, ud = lower_circuit(W.qft_circuit(n))
        else:
            ops, ud = random_ops(rng, n, 40)
        st = W.normal_state(n, seed + 100 + i, dtype)
        jobs.append((n, dtype, ops, ud, st))
    # Grover circuits above one tile: their phase flips and diffusers run as
    # batched launches; several of them share rounds
    for i, (n, dtype) in enumerate(((14, np.float64), (15, np.float32), (14, np.float32), (13, np.float64))):
        qc, _ = W.grover_search(n, seed + 50 + i)
        ops, ud = lower_circuit(qc)
        jobs.append((n, dtype, ops, ud, W.normal_state(n, seed + 200 + i, dtype)))
    return jobs


def run_alone(job):
    n, dtype, ops, ud, st = job
    g = GpuState.from_host(st)
    p = Plan(n, dtype, ops, ud, _lib.opts())
    p.execute(g)
    out = g.download()
    g.free()
    p.free()
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_alone(gpu_lib, graph):
    jobs = mixed_jobs(7)
    alone = [run_alone(j) for j in jobs]
    states = [GpuState.from_host(j[4]) for j in jobs]
    plans = [Plan(j[0], j[1], j[2], j[3], _lib.opts()) for j in jobs]
    b = Batch(states, plans, graph=graph)
    assert b.stats()["rounds"] >= 1
    b.launch()
    b.synchronize()
    for st, ref in zip(states, alone):
        out = st.download()
        assert np.array_equal(out.re.view(np.uint8), ref.re.view(np.uint8))
        assert np.array_equal(out.im.view(np.uint8), ref.im.view(np.uint8))
    b.free()
    for s in states:
        s.free()


def test_graph_replays_match_eager(gpu_lib):
    jobs = mixed_jobs(11, count=8)
    init = [j[4] for j in jobs]
    eager_states = [GpuState.from_host(s) for s in init]
    graph_states = [GpuState.from_host(s) for s in init]
    plans_e = [Plan(j[0], j[1], j[2], j[3], _lib.opts()) for j in jobs]
    plans_g = [Plan(j[0], j[1], j[2], j[3], _lib.opts()) for j in jobs]
    be = Batch(eager_states, plans_e, graph=False)
    bg = Batch(graph_states, plans_g, graph=True)
    for rep in range(3):
        be.launch()
        be.synchronize()
        # eager work and a synchronisation between graph replays
        for s in graph_states:
            s.norm2()
        bg.launch()
        bg.synchronize()
        for se, sg in zip(eager_states, graph_states):
            a, c = se.download(), sg.download()
            assert np.array_equal(a.re, c.re) and np.array_equal(a.im, c.im), rep
    be.free()
    bg.free()


def test_graph_of_more_than_2048_launches_is_refused(gpu_lib):
    """A batch graph is limited to 2048 launches (BFLY_EUNSUPPORTED beyond,
    before any capture); the same batch without a graph is accepted."""
    n = 10
    ops, ud = lower_circuit(W.random_layered(n, 240, 4))
    plan = Plan(n, np.float32, ops, ud, _lib.opts(tile_bits=8, max_gates_per_pass=1))
    assert plan.stats()["launches"] > 2048
    g = GpuState(n, np.float32)
    with pytest.raises(_lib.BflyError, match="launches"):
        Batch([g], [plan], graph=True)
    b = Batch([g], [plan], graph=False)
    assert b.stats()["launches"] > 2048
    b.free()
    plan.free()
    g.free()
