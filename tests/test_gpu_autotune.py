"""GPU: the autotuner.  The tuned options are one of the candidates, never
slower than the base options by its own measurement, and the state they
produce equals the oracle's; the cache tunes a class once; the engine and
the batch runner use the tuned options and still match the oracle and each
other."""
import numpy as np
import pytest

from code_mojo_amd import _lib, autotune, serve
from code_mojo_amd.state import GpuState, Plan
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_state(job):
    re = np.zeros(1 << job.n)
    im = np.zeros(1 << job.n)
    re[0] = 1.0
    O.execute(job.n, re, im, job.ops, job.udata)
    return re + 1j * im


def run(job, opts):
    g = GpuState(job.n, job.dtype)
    p = Plan(job.n, job.dtype, job.ops, job.udata, opts)
    p.execute(g)
    s = g.download()
    p.free()
    g.free()
    return s.re + 1j * s.im


@pytest.mark.parametrize("make,dtype,tol", [
    (lambda: serve.layered_job(12, 8, 3, np.float64), np.float64, 1e-12),
    (lambda: serve.qft_job(12, 4, np.float64), np.float64, 1e-12),
    (lambda: serve.grover_job(11, 5, np.float32), np.float32, 2e-5),
])
def test_tuned_options_are_a_candidate_and_exact(gpu_lib, make, dtype, tol):
    job = make()
    job.dtype = dtype
    job.init_kind = _lib.INIT_ZERO
    r = autotune.tune(job.n, job.dtype, job.ops, job.udata, repeats=2)
    sigs = {(o.mode, o.tile_bits, o.max_gates_per_pass) for o in autotune.candidates(job.n, job.dtype)}
    assert (r.opts.mode, r.opts.tile_bits, r.opts.max_gates_per_pass) in sigs
    assert 0.0 < r.best_ms <= r.base_ms
    got = run(job, r.opts)
    assert np.abs(got - oracle_state(job)).max() <= tol


def test_cache_tunes_each_class_once(gpu_lib):
    t = autotune.Autotuner(repeats=1)
    a = serve.layered_job(10, 4, 1, np.float32)
    b = serve.layered_job(10, 4, 2, np.float32)   # same class (family, qubits, dtype)
    c = serve.layered_job(11, 4, 3, np.float32)
    t.opts_for(a)
    t.opts_for(b)
    t.opts_for(c)
    assert len(t.cache) == 2
    assert set(t.summary()) == {"layered/10/float32", "layered/11/float32"}
    wide = serve.layered_job(12, 2, 4, np.float32)
    t.max_qubits = 11
    assert t.opts_for(wide) is t.base and len(t.cache) == 2


def test_engine_and_batch_use_tuned_options(gpu_lib):
    tuner = autotune.Autotuner(repeats=1)
    jobs = [serve.layered_job(10 + i % 3, 5, 40 + i, np.float32) for i in range(6)]
    jobs += [serve.grover_job(9 + i % 2, 60 + i, np.float32) for i in range(4)]
    with serve.Engine(tuner=tuner) as eng:
        res = serve.run_stream(jobs, None, eng, keep_results=True)
    assert res["completed"] == len(jobs), res["failed"]
    assert len(tuner.cache) == 5
    for job, st in zip(jobs, res["results"]):
        ref = oracle_state(job)
        assert np.abs((st.re + 1j * st.im) - ref).max() <= 2e-5
    out = serve.bench_batching(jobs, graph=True, repeats=1, check=True, tuner=tuner)
    assert out["max_amp_err_vs_seq"] == 0.0          # batch and one-by-one runs agree bit for bit
    assert out["min_fidelity_vs_c128"] > 1 - 1e-9
    assert set(out["tuned"]) == set(tuner.summary())
