"""Host side of the serving workloads: request builders, stream mix,
arrival process (no GPU)."""
import numpy as np

from code_mojo_amd import lower_circuit, serve
from code_mojo_amd import workloads as W


def test_grover_ops_equal_full_lowering():
    for n in (3, 5, 8):
        for seed in (1, 2):
            qc, marked = W.grover_search(n, seed)
            ops, ud = lower_circuit(qc)
            job = serve.grover_job(n, seed, np.float64)
            assert job.meta["marked"] == marked
            assert np.array_equal(job.ops, ops) and ud.size == job.udata.size == 0


def test_serving_stream_mix_and_determinism():
    spec = serve.StreamSpec(count=200, lo=8, hi=14, layered_max=12, qft_max=14, grover_max=9)
    a = serve.serving_jobs(spec)
    b = serve.serving_jobs(spec)
    assert [j.name for j in a] == [j.name for j in b]
    assert all(np.array_equal(x.ops, y.ops) for x, y in zip(a, b))
    fam = [j.name.rstrip("0123456789x") for j in a]
    share = {f: fam.count(f) / len(a) for f in set(fam)}
    assert abs(share["layered"] - 0.4) < 0.1 and abs(share["qft"] - 0.3) < 0.1
    assert abs(share["grover"] - 0.2) < 0.1 and abs(share["ae"] - 0.1) < 0.07
    assert all(j.n <= 12 for j in a if j.name.startswith("layered"))
    assert all(j.n <= 9 for j in a if j.name.startswith("grover"))
    dts = [j.dtype for j in a]
    assert 0.35 < dts.count(np.float64) / len(a) < 0.65
    big = serve.serving_jobs(serve.StreamSpec(count=10, big=True, big_n=20))
    assert len(big) == 11 and any(j.n == 20 and j.name == "qft20" for j in big)


def test_poisson_arrivals_rate():
    t = serve.poisson_arrivals(20000, 250.0, 3)
    assert np.all(np.diff(t) > 0)
    assert abs(20000 / t[-1] - 250.0) < 10.0


def test_batching_jobs_shape():
    jobs = serve.batching_jobs(64, 1, lo=6, hi=12, grover_max=10, layers=4)
    assert len(jobs) == 64
    assert all(6 <= j.n <= 12 for j in jobs)
    assert all(j.name.startswith("grover") for j in jobs[::2])
    assert all(j.name.startswith("layered") for j in jobs[1::2])
    assert all(j.dtype == np.float32 for j in jobs)
