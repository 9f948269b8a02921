"""Serving loop and batch driver around the native engine (bfly_engine).

The engine (code_mojo_amd/csrc/engine.hip) admits circuits against an HBM
budget, keeps their states resident and advances every running circuit
together, fusing the tile passes of a round into one batched launch per
tile size.  This module builds the seeded request streams of the serving
and batching workloads, submits them (at Poisson arrival times for the
serving stream), and reports circuits/s and latency percentiles.  The
reference has no serving layer; its per-call `execute()` is the
"one circuit per launch" baseline measured by `run_sequential`.
"""
from __future__ import annotations

import ctypes as ct
import math
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib, algos, workloads as W
from ._lib import check
from .circuit import QuantumCircuit
from .lowering import lower_circuit
from .state import Batch, GpuState, Plan, QuantumState, dtype_code


@dataclass
class Job:
    """One circuit to serve: lowered ops + initial state recipe."""

    name: str
    n: int
    dtype: type
    ops: np.ndarray
    udata: np.ndarray
    init_kind: int = _lib.INIT_ZERO
    init_seed: int = 0
    init: Optional[QuantumState] = None
    meta: Dict = field(default_factory=dict)


def default_exec_opts() -> _lib.ExecOpts:
    """Served circuits compile in AUTO mode: the planner keeps the faster of
    the tile-pass and the dense-block schedule (QFT-type ladders run as
    MFMA blocks, layered circuits as tile passes)."""
    return _lib.opts(mode=_lib.MODE_AUTO)


def qasm_job(name: str, text: str, dtype=np.float64) -> Job:
    """A job from OpenQASM 3 text, parsed and lowered by the library's C++
    front end (no Python circuit objects on the serving path)."""
    n, ops, udata = _lib.qasm3_lower(text)
    return Job(name=name, n=n, dtype=dtype, ops=ops, udata=udata)


class Engine:
    """Python handle of a bfly_engine."""

    def __init__(self, budget_bytes: int = 0, max_inflight_rounds: int = 0, large_qubits: int = 0,
                 max_round_requests: int = 0, planner_threads: int = 0, tuner=None):
        """`tuner` (an autotune.Autotuner): requests submitted without options
        compile with the options tuned for their class instead of AUTO."""
        self.tuner = tuner
        lib = _lib.load()
        opts = _lib.EngineOpts(budget_bytes, max_inflight_rounds, large_qubits, max_round_requests, 0,
                               planner_threads, 0)
        h = ct.c_void_p()
        check(lib.bfly_engine_create(ct.byref(h), ct.byref(opts)))
        self._h = h
        self._keep: Dict[int, Tuple] = {}

    def submit(self, job: Job, exec_opts: Optional[_lib.ExecOpts] = None) -> int:
        rq = _lib.RequestDesc()
        rq.num_qubits = job.n
        rq.dtype = dtype_code(job.dtype)
        rq.ops = _lib.ops_ptr(job.ops)
        rq.nops = job.ops.size
        rq.udata = _lib.udata_ptr(job.udata)
        rq.udata_len = job.udata.size // 2
        rq.init_kind = job.init_kind
        rq.init_seed = job.init_seed
        if job.init_kind == _lib.INIT_HOST:
            rq.init_re = job.init.re.ctypes.data
            rq.init_im = job.init.im.ctypes.data
        if exec_opts is None:
            exec_opts = self.tuner.opts_for(job) if self.tuner is not None else default_exec_opts()
        rq.exec = exec_opts
        rid = ct.c_uint64()
        check(_lib.load().bfly_engine_submit(self._h, ct.byref(rq), ct.byref(rid)))
        return rid.value

    def status(self, rid: int) -> Tuple[int, float, float, float]:
        st, a, b, c = ct.c_int32(), ct.c_double(), ct.c_double(), ct.c_double()
        check(_lib.load().bfly_engine_status(self._h, rid, ct.byref(st), ct.byref(a), ct.byref(b), ct.byref(c)))
        return st.value, a.value, b.value, c.value

    def wait(self, rid: int, timeout: float = -1.0) -> None:
        check(_lib.load().bfly_engine_wait(self._h, rid, timeout))

    def fetch(self, rid: int, n: int, dtype) -> QuantumState:
        out = QuantumState(n, dtype)
        check(_lib.load().bfly_engine_fetch(self._h, rid, out.re.ctypes.data, out.im.ctypes.data, out.size()))
        return out

    def fetch_range(self, rid: int, first: int, count: int, dtype) -> np.ndarray:
        """Amplitudes [first, first + count) of a finished request, as complex
        numbers (the request stays until fetched or released)."""
        re = np.empty(count, dtype=dtype)
        im = np.empty(count, dtype=dtype)
        check(_lib.load().bfly_engine_fetch_range(self._h, rid, first, count, re.ctypes.data, im.ctypes.data))
        return re.astype(np.float64) + 1j * im.astype(np.float64)

    def release(self, rid: int) -> None:
        check(_lib.load().bfly_engine_release(self._h, rid))

    def cancel(self, rid: int) -> None:
        check(_lib.load().bfly_engine_cancel(self._h, rid))

    def error(self, rid: int) -> str:
        buf = ct.create_string_buffer(512)
        check(_lib.load().bfly_engine_error(self._h, rid, buf, 512))
        return buf.value.decode()

    def stats(self) -> dict:
        a, c, r, l = ct.c_int64(), ct.c_int64(), ct.c_int64(), ct.c_int64()
        u, p = ct.c_uint64(), ct.c_uint64()
        check(_lib.load().bfly_engine_stats(self._h, ct.byref(a), ct.byref(c), ct.byref(r), ct.byref(l),
                                            ct.byref(u), ct.byref(p)))
        return {"admitted": a.value, "completed": c.value, "rounds": r.value, "launches": l.value,
                "bytes_in_use": u.value, "peak_bytes": p.value}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _lib.load().bfly_engine_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -- workloads ----------------------------------------------------------------
def layered_job(n: int, depth: int, seed: int, dtype) -> Job:
    ops, ud = lower_circuit(W.random_layered(n, depth, seed))
    return Job(f"layered{n}x{depth}", n, dtype, ops, ud)


def qft_job(n: int, seed: int, dtype) -> Job:
    ops, ud = lower_circuit(W.qft_circuit(n))
    return Job(f"qft{n}", n, dtype, ops, ud, _lib.INIT_RANDOM, seed)


def grover_ops(n: int, marked: int) -> Tuple[np.ndarray, np.ndarray]:
    """Lowered ops of `algos.grover_search_circuit(n, marked)`, built from one
    lowered iterate repeated (lowering is gate by gate, so the result is
    identical to lowering the whole circuit, in a fraction of the time)."""
    iters = int(math.floor(math.pi / 4 * math.sqrt(2 ** n)))
    head = QuantumCircuit(n)
    for i in range(n):
        head.h(i)
    h_ops, h_ud = lower_circuit(head)
    step = algos.grover_iterate_circuit(algos.phase_oracle_match(n, [marked]), n)
    s_ops, s_ud = lower_circuit(step)
    assert h_ud.size == 0 and s_ud.size == 0
    return np.concatenate([h_ops, np.tile(s_ops, iters)]), h_ud


def grover_job(n: int, seed: int, dtype) -> Job:
    marked = int(W.splitmix64(seed, 1)[0] % np.uint64(1 << n))
    ops, ud = grover_ops(n, marked)
    return Job(f"grover{n}", n, dtype, ops, ud, meta={"marked": marked})


def ae_job(seed: int, dtype, eval_qubits: int = 8) -> Job:
    qc, a = W.amplitude_estimation(seed, eval_qubits)
    ops, ud = lower_circuit(qc)
    return Job(f"ae{qc.num_qubits}", qc.num_qubits, dtype, ops, ud, meta={"a": a})


def batching_jobs(count: int, seed: int, lo: int = 10, hi: int = 22, dtype=np.float32,
                  grover_max: int = 22, layers: int = 20) -> List[Job]:
    """Batching workload: `count` circuits, qubit counts log-uniform in
    [lo, hi]; half Grover (one seeded mark, floor(pi/4 sqrt(2^n))
    iterations), half `layers`-layer random U3+CZ brick circuits."""
    ns = W.loguniform_qubits(seed, count, lo, hi)
    jobs = []
    for i, n in enumerate(ns):
        if i % 2 == 0:
            jobs.append(grover_job(min(n, grover_max), seed * 7919 + i, dtype))
        else:
            jobs.append(layered_job(n, layers, seed * 104729 + i, dtype))
    return jobs


@dataclass
class StreamSpec:
    """Serving stream: Poisson arrivals at `rate` per second; mix 40 %
    layered (depth 20-60), 30 % QFT on seeded random states, 20 % Grover,
    10 % amplitude estimation (8 evaluation qubits); half complex128, half
    complex64.  Qubit counts are drawn uniformly from [lo, hi] and capped
    per family so that one request stays within a few seconds of GPU work
    (Grover <= grover_max, layered <= layered_max); `big` adds one
    `big_n`-qubit complex128 QFT to the stream."""

    count: int = 200
    rate: float = 200.0
    lo: int = 12
    hi: int = 32
    grover_max: int = 16
    layered_max: int = 26
    qft_max: int = 30
    seed: int = 5
    big: bool = False
    big_n: int = 33


def serving_jobs(spec: StreamSpec) -> List[Job]:
    u = W.uniforms(spec.seed, 8 * spec.count)
    jobs: List[Job] = []
    for i in range(spec.count):
        r, rn, rd, rdepth = u[8 * i], u[8 * i + 1], u[8 * i + 2], u[8 * i + 3]
        n = spec.lo + int(rn * (spec.hi - spec.lo + 1))
        dtype = np.float64 if rd < 0.5 else np.float32
        seed = spec.seed * 1000003 + i
        if r < 0.4:
            jobs.append(layered_job(min(n, spec.layered_max), 20 + int(rdepth * 41), seed, dtype))
        elif r < 0.7:
            jobs.append(qft_job(min(n, spec.qft_max), seed, dtype))
        elif r < 0.9:
            jobs.append(grover_job(min(n, spec.grover_max), seed, dtype))
        else:
            jobs.append(ae_job(seed, dtype))
    if spec.big:
        jobs.insert(len(jobs) // 2, qft_job(spec.big_n, spec.seed, np.float64))
    return jobs


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    u = W.uniforms(seed + 17, count)
    gaps = -np.log(1.0 - u) / rate
    return np.cumsum(gaps)


def percentile(xs: Sequence[float], q: float) -> float:
    if not xs:
        return float("nan")
    return float(np.percentile(np.asarray(xs), q))


def run_stream(jobs: List[Job], arrivals: Optional[np.ndarray], engine: Engine, timeout: float = 600.0,
               keep_results: bool = False) -> dict:
    """Submit jobs (at arrival times if given, else all at once), wait for
    all, and report throughput and latency (submit -> done)."""
    t0 = time.perf_counter()
    ids = []
    for i, job in enumerate(jobs):
        if arrivals is not None:
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        ids.append(engine.submit(job))
    lat, failed = [], []
    results = []
    for rid, job in zip(ids, jobs):
        engine.wait(rid, timeout)
        st, ts, ta, td = engine.status(rid)
        if st != _lib.REQ_DONE:
            failed.append((job.name, st, engine.error(rid) if st == _lib.REQ_FAILED else ""))
            engine.release(rid)
            continue
        lat.append(td - ts)
        if keep_results:
            results.append(engine.fetch(rid, job.n, job.dtype))
        else:
            engine.release(rid)
    wall = time.perf_counter() - t0
    out = {"circuits": len(jobs), "completed": len(lat), "failed": failed, "wall_s": wall,
           "circuits_per_s": len(lat) / wall if wall > 0 else float("nan"),
           "p50_ms": percentile(lat, 50) * 1e3, "p99_ms": percentile(lat, 99) * 1e3,
           "max_ms": (max(lat) * 1e3) if lat else float("nan"), "engine": engine.stats()}
    if keep_results:
        out["results"] = results
    return out


def run_sequential(jobs: List[Job], arrivals: Optional[np.ndarray] = None, reuse_states: bool = True) -> dict:
    """One circuit per launch sequence: each circuit planned and executed
    alone, one after the other (the reference's execute() call pattern).
    With `reuse_states` one device state per (qubits, dtype) is kept and
    re-initialised for every circuit, so the comparison with the engine
    (which pools states) is not about allocation cost."""
    t0 = time.perf_counter()
    lat = []
    cache: Dict[Tuple[int, type], GpuState] = {}
    for i, job in enumerate(jobs):
        if arrivals is not None:
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        ts = time.perf_counter()
        key = (job.n, job.dtype)
        g = cache.get(key) if reuse_states else None
        if g is None:
            g = GpuState(job.n, job.dtype)
            if reuse_states:
                cache[key] = g
        elif job.init_kind == _lib.INIT_ZERO:
            g.reset()
        if job.init_kind == _lib.INIT_RANDOM:
            g.fill_random(job.init_seed)
        elif job.init_kind == _lib.INIT_HOST:
            g.upload(job.init)
        p = Plan(job.n, job.dtype, job.ops, job.udata, default_exec_opts())
        p.execute(g)
        g.synchronize()
        t_done = time.perf_counter()
        ref = ts if arrivals is None else t0 + arrivals[i]
        lat.append(t_done - ref)
        p.free()
        if not reuse_states:
            g.free()
    wall = time.perf_counter() - t0
    for g in cache.values():
        g.free()
    return {"circuits": len(jobs), "wall_s": wall, "circuits_per_s": len(jobs) / wall,
            "p50_ms": percentile(lat, 50) * 1e3, "p99_ms": percentile(lat, 99) * 1e3}


# -- benchmarks -----------------------------------------------------------------
def _init_state(job: Job) -> GpuState:
    g = GpuState(job.n, job.dtype)
    if job.init_kind == _lib.INIT_RANDOM:
        g.fill_random(job.init_seed)
    elif job.init_kind == _lib.INIT_HOST:
        g.upload(job.init)
    return g


def _reinit(g: GpuState, job: Job) -> None:
    if job.init_kind == _lib.INIT_RANDOM:
        g.fill_random(job.init_seed)
    elif job.init_kind == _lib.INIT_HOST:
        g.upload(job.init)
    else:
        g.reset()


def fidelity(a: QuantumState, b: QuantumState) -> Tuple[float, float]:
    """(L2 error ||a - b||, fidelity |<a|b>|^2 / (<a|a><b|b>)) in FP64."""
    ar, ai = a.re.astype(np.float64), a.im.astype(np.float64)
    br, bi = b.re.astype(np.float64), b.im.astype(np.float64)
    l2 = math.sqrt(float(np.sum((ar - br) ** 2) + np.sum((ai - bi) ** 2)))
    ip_re = float(np.dot(ar, br) + np.dot(ai, bi))
    ip_im = float(np.dot(ar, bi) - np.dot(ai, br))
    na = float(np.dot(ar, ar) + np.dot(ai, ai))
    nb = float(np.dot(br, br) + np.dot(bi, bi))
    return l2, (ip_re * ip_re + ip_im * ip_im) / (na * nb)


def bench_batching(jobs: List[Job], graph: bool = True, repeats: int = 1, check: bool = True,
                   exec_opts: Optional[_lib.ExecOpts] = None, log=None, check_every: int = 1,
                   tuner=None) -> dict:
    """Run `jobs` (i) as one batch: every round's tile passes of all circuits
    in one persistent launch per tile size, all rounds captured in one
    hipGraph, and (ii) one circuit after another, each circuit's launches
    alone (the per-call pattern of the reference's execute()).  States start
    from their recipe before every timed run; plans are compiled outside the
    timed region for both.  With `check`, every final state is compared with
    a complex128 run of the same circuit (every `check_every`-th circuit
    against its complex128 run; all of them against the one-by-one run).
    With `tuner` (an autotune.Autotuner) each circuit compiles with the
    options tuned for its class (tile shape, gates per pass, schedule mode),
    tuned before anything is timed; both runs use the same plans."""
    eo = exec_opts or _lib.opts()
    say = log or (lambda m: None)
    t0 = time.perf_counter()
    job_opts = [tuner.opts_for(j) for j in jobs] if tuner is not None else [eo] * len(jobs)
    if tuner is not None:
        say(f"batching: tuned {len(tuner.cache)} circuit classes in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    states = [_init_state(j) for j in jobs]
    plans = [Plan(j.n, j.dtype, j.ops, j.udata, o) for j, o in zip(jobs, job_opts)]
    say(f"batching: {len(jobs)} states + plans in {time.perf_counter() - t0:.1f}s")
    passes = sum(p.stats()["launches"] for p in plans)
    algo_bytes = sum(p.stats()["launches"] * 2 * (1 << j.n) * np.dtype(j.dtype).itemsize * 2
                     for p, j in zip(plans, jobs))
    batch = Batch(states, plans, graph=graph)
    batch.launch()              # first launch captures the graph
    batch.synchronize()
    results = [g.download() for g in states] if check else None
    t_batch = []
    for _ in range(repeats):
        for g, j in zip(states, jobs):
            _reinit(g, j)
        for g in states:
            g.synchronize()
        ts = time.perf_counter()
        batch.launch()
        batch.synchronize()
        t_batch.append(time.perf_counter() - ts)
    bstats = batch.stats()
    batch.free()
    say(f"batching: batch {min(t_batch):.3f}s ({bstats})")
    # one circuit per launch sequence
    for g, j in zip(states, jobs):
        _reinit(g, j)
    for g in states:
        g.synchronize()
    plans[0].execute(states[0])   # warm the kernels
    states[0].synchronize()
    _reinit(states[0], jobs[0])
    states[0].synchronize()
    ts = time.perf_counter()
    for g, p in zip(states, plans):
        p.execute(g)
        g.synchronize()
    t_seq = time.perf_counter() - ts
    say(f"batching: sequential {t_seq:.3f}s")
    worst = {"max_amp_err_vs_seq": 0.0}
    if check:
        for g, r in zip(states, results):
            s = g.download()
            worst["max_amp_err_vs_seq"] = max(worst["max_amp_err_vs_seq"],
                                              float(np.max(np.abs(s.re - r.re))), float(np.max(np.abs(s.im - r.im))))
    for p in plans:
        p.free()
    for g in states:
        g.free()
    out = {"circuits": len(jobs), "batch_s": min(t_batch), "sequential_s": t_seq,
           "circuits_per_s": len(jobs) / min(t_batch), "sequential_circuits_per_s": len(jobs) / t_seq,
           "speedup": t_seq / min(t_batch), "passes": passes, "algorithmic_bytes": algo_bytes,
           "rounds": bstats["rounds"], "launches": bstats["launches"], "graph": graph}
    if tuner is not None:
        out["tuned"] = tuner.summary()
        out["tune_s"] = tuner.seconds()
    if check:
        l2s, infid = [], []
        for j, r in list(zip(jobs, results))[::check_every]:
            g = GpuState(j.n, np.float64)
            p = Plan(j.n, np.float64, j.ops, j.udata, _lib.opts())
            _reinit(g, j)
            p.execute(g)
            ref = g.download()
            p.free()
            g.free()
            l2, f = fidelity(r, ref)
            l2s.append(l2)
            infid.append(1.0 - f)
        out.update({"max_l2_vs_c128": max(l2s), "min_fidelity_vs_c128": 1.0 - max(infid),
                    "checked_vs_c128": len(l2s), **worst})
        say(f"batching: checked {len(l2s)} results vs complex128 runs")
    return out


def bench_serving(jobs: List[Job], arrivals: Optional[np.ndarray], engine_opts: Optional[dict] = None,
                  sequential: bool = True, log=None) -> dict:
    """The same request list through the engine and through one circuit per
    launch sequence; `arrivals` None means every request at t=0 (saturated
    throughput), else the Poisson arrival times."""
    say = log or (lambda m: None)
    with Engine(**(engine_opts or {})) as eng:
        # warm-up: one small request of each dtype
        for dt in (np.float32, np.float64):
            rid = eng.submit(layered_job(10, 2, 1, dt))
            eng.wait(rid, 60)
            eng.release(rid)
        eng_res = run_stream(jobs, arrivals, eng)
    say(f"serving: engine {eng_res['circuits_per_s']:.1f} circuits/s p50 {eng_res['p50_ms']:.2f} ms "
        f"p99 {eng_res['p99_ms']:.2f} ms failed {len(eng_res['failed'])}")
    out = {"engine": {k: v for k, v in eng_res.items() if k != "results"}}
    if sequential:
        run_sequential(jobs[:2])  # warm-up
        seq = run_sequential(jobs, arrivals)
        say(f"serving: sequential {seq['circuits_per_s']:.1f} circuits/s p50 {seq['p50_ms']:.2f} ms "
            f"p99 {seq['p99_ms']:.2f} ms")
        out["sequential"] = seq
        out["speedup"] = eng_res["circuits_per_s"] / seq["circuits_per_s"]
    return out
