"""Benchmark of the GPU statevector strategy on one MI355X.

Default workload (the headline metric's configuration): a 30-qubit
complex128 random layered circuit, 40 layers of U3 on every qubit + CZ brick
(offset alternating 0/1) + one CX on a random pair, from |0..0>; the state
(16 GiB) stays resident in HBM.  A step is one execution of the whole
circuit's compiled pass schedule.

Printed: ONE JSON line with
  value        state GB/s = passes * 2 * 2^n * 16 B / wall time of a step
  roofline     tile_pass_kernel: algorithmic bytes per launch (2 * 2^n * b)
               / mean launch time from HIP events on the state's stream
               during the timed region, against 8 TB/s
  cpu_baseline the reference's fastest CPU strategy (GRID_PARALLEL_FUSED,
               C/OpenMP restatement) on a bounded sample of the same workload
               on this host
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_info() -> dict:
    """Host description for the CPU baseline: logical CPUs, the CPU model."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(n: int, layers: int, seed: int, threads: int) -> dict:
    """GRID_PARALLEL_FUSED restatement on the host cores: the first `layers`
    layers of the headline circuit (same generator and seed) at `n` qubits,
    with the reference's default grid (slack 3: 8 rows) and with slack
    log2(threads) (rows for every thread, the reference's own knob)."""
    from code_mojo_amd import lower_circuit
    from code_mojo_amd import workloads as W
    from oracle import oracle as O

    O.load()
    qc = W.random_layered(n, layers, seed)
    ops, _ = lower_circuit(qc)
    re = np.zeros(1 << n)
    im = np.zeros(1 << n)
    re[:] = 0.0                 # touch every page before anything is timed
    im[:] = 0.0
    re[0] = 1
    res = {}
    for label, slack in (("faithful_slack3", 3), ("tuned_slack_log2_threads", max(3, int(np.log2(max(threads, 1)))))):
        t0 = time.perf_counter()
        O.grid_fused_execute(n, re, im, ops, slack=slack, threads=threads)
        res[label] = time.perf_counter() - t0
    best = min(res.values())
    state_bytes = 2 * (1 << n) * 16
    return {
        # gate-equivalent state traffic: one state read + write per gate (the
        # grid strategy walks every row once per gate)
        "value": round(len(ops) * state_bytes / best / 1e9, 3),
        "unit": "GB/s (gate-equivalent: one state read and write per gate)",
        "cores": threads,
        "kind": "port",
        "sample": f"GRID_PARALLEL_FUSED C/OpenMP restatement on the headline circuit: {n} qubits complex128, "
                  f"the first {layers} of its 40 layers ({len(ops)} gates), best of the reference's default "
                  f"grid (slack 3) and slack log2(threads), one timed run each on a state whose pages are "
                  f"already touched; {threads} OpenMP threads; state allocation and circuit build excluded",
        "gate_layers_per_s": round(layers / best, 5),
        "seconds": {k: round(v, 3) for k, v in res.items()},
        **cpu_info(),
    }


def bench_batching(args) -> dict:
    """4096 independent complex64 circuits, qubit counts log-uniform in
    [10, 22]: half Grover search (one seeded mark), half 20-layer random
    U3 + CZ brick circuits.  Timed: one hipGraph launch running every round
    of the batch (states reset between timed runs), against the same
    circuits executed one after another.  Every circuit compiles with the
    options the autotuner picked for its class (family, qubits; tuned on the
    GPU before anything is timed; --no-autotune for the defaults).  Every 8th
    final state is compared with a complex128 run of its circuit."""
    from code_mojo_amd import autotune, serve

    count = args.circuits or 4096
    t0 = time.perf_counter()
    jobs = serve.batching_jobs(count, args.seed)
    log(f"[bench] {count} batching circuits built in {time.perf_counter() - t0:.1f}s")
    tuner = None if args.no_autotune else autotune.Autotuner()
    res = serve.bench_batching(jobs, graph=True, repeats=max(args.steps, 1), check=True, log=log, check_every=8,
                               tuner=tuner)
    secs = res["batch_s"]
    out = {
        "metric": "circuits/s, batched persistent launches (complex64, 10-22 qubits)",
        "value": round(res["circuits_per_s"], 2), "unit": "circuits/s", "n_gpus": 1,
        "steps": max(args.steps, 1), "warmup": 1, "ms_per_step": round(secs * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Grover and random layered circuits)",
        "config": {"workload": "batching_grover_layered", "circuits": count, "qubits": "10-22 log-uniform",
                   "seed": args.seed, "graph": True, "tile_passes": res["passes"], "rounds": res["rounds"],
                   "launches": res["launches"], "autotuned": tuner is not None},
        "tuned": res.get("tuned"), "tune_s": round(res["tune_s"], 2) if "tune_s" in res else None,
        "one_circuit_per_launch_circuits_per_s": round(res["sequential_circuits_per_s"], 2),
        "speedup_vs_one_circuit_per_launch": round(res["speedup"], 3),
        "max_l2_vs_c128": res.get("max_l2_vs_c128"), "min_fidelity_vs_c128": res.get("min_fidelity_vs_c128"),
        "checked_vs_c128": res.get("checked_vs_c128"),
        "max_amp_err_batch_vs_alone": res.get("max_amp_err_vs_seq"),
        "roofline": {"bound": "hbm", "achieved": round(res["algorithmic_bytes"] / secs / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(res["algorithmic_bytes"] / secs / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batch (every launch of the graph)"},
        "cpu_baseline": None,
    }
    return out


def bench_serving(args) -> dict:
    """A seeded Poisson stream (40 % layered, 30 % QFT on random states,
    20 % Grover, 10 % amplitude estimation; half complex128, half complex64)
    through the serving engine, and the same stream one circuit at a time.
    Reported: throughput with every request submitted at once (capacity), and
    p50 / p99 latency at the stated arrival rate."""
    from code_mojo_amd import serve

    spec = serve.StreamSpec(count=args.circuits or 200, rate=args.rate, seed=args.seed % 1000 + 5, big=args.big)
    t0 = time.perf_counter()
    jobs = serve.serving_jobs(spec)
    log(f"[bench] {len(jobs)} serving requests built in {time.perf_counter() - t0:.1f}s")
    sat = serve.bench_serving(jobs, None, log=log)
    if spec.rate <= 0:
        # a sustainable arrival rate: 70 % of the capacity just measured, so
        # the latency percentiles describe service, not an ever-growing queue
        spec.rate = round(0.7 * sat["engine"]["circuits_per_s"], 1)
    arrivals = serve.poisson_arrivals(len(jobs), spec.rate, spec.seed)
    live = serve.bench_serving(jobs, arrivals, log=log)
    eng = sat["engine"]
    # the same mix restricted to 12-20 qubits: every request is launch-bound
    # alone, which is where batching across requests pays
    small_spec = serve.StreamSpec(count=spec.count, rate=args.rate, seed=spec.seed, lo=12, hi=20)
    small_jobs = serve.serving_jobs(small_spec)
    small = serve.bench_serving(small_jobs, None, log=log)
    if small_spec.rate <= 0:
        small_spec.rate = round(0.7 * small["engine"]["circuits_per_s"], 1)
    small_live = serve.bench_serving(small_jobs, serve.poisson_arrivals(len(small_jobs), small_spec.rate, spec.seed),
                                     log=log)
    out = {
        "metric": "circuits/s through the serving engine (saturated), p99 latency at the stated rate",
        "value": round(eng["circuits_per_s"], 2), "unit": "circuits/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(eng["wall_s"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic (seeded request stream)",
        "config": {"workload": "serving_stream", "requests": len(jobs), "rate_per_s": spec.rate,
                   "qubits": f"{spec.lo}-{spec.hi}", "caps": {"grover": spec.grover_max,
                                                              "layered": spec.layered_max, "qft": spec.qft_max},
                   "big_qft_qubits": spec.big_n if spec.big else None, "seed": spec.seed},
        "one_circuit_per_launch_circuits_per_s": round(sat["sequential"]["circuits_per_s"], 2),
        "speedup_vs_one_circuit_per_launch": round(sat["speedup"], 3),
        "saturated": {"p50_ms": round(eng["p50_ms"], 3), "p99_ms": round(eng["p99_ms"], 3),
                      "failed": len(eng["failed"])},
        "at_rate": {"rate_per_s": spec.rate,
                    "engine": {k: round(live["engine"][k], 3) for k in ("circuits_per_s", "p50_ms", "p99_ms", "max_ms")},
                    "one_circuit_per_launch": {k: round(live["sequential"][k], 3)
                                               for k in ("circuits_per_s", "p50_ms", "p99_ms")}},
        "engine_stats": eng["engine"],
        "small_requests_12_20q": {
            "circuits_per_s": round(small["engine"]["circuits_per_s"], 2),
            "one_circuit_per_launch_circuits_per_s": round(small["sequential"]["circuits_per_s"], 2),
            "speedup_vs_one_circuit_per_launch": round(small["speedup"], 3),
            "saturated_p99_ms": round(small["engine"]["p99_ms"], 3),
            "at_rate": {"rate_per_s": small_spec.rate,
                        "engine": {k: round(small_live["engine"][k], 3) for k in ("circuits_per_s", "p50_ms", "p99_ms")},
                        "one_circuit_per_launch": {k: round(small_live["sequential"][k], 3)
                                                   for k in ("circuits_per_s", "p50_ms", "p99_ms")}}},
        "roofline": None,
        "cpu_baseline": None,
    }
    return out


def bench_precision(args) -> dict:
    """Value-encoding circuit (H on every qubit, exact dyadic phases for a
    seeded n-bit integer, inverse QFT, bit reversal) in complex64 BLOCK mode
    (5-qubit 32 x 32 blocks on FP32 MFMA, FP64 phase generation, compensated
    norm check and renormalisation every 8 blocks), against the plain
    complex128 GPU path (one launch per gate) and the fused complex128 tile
    path.  L2 error and fidelity of the complex64 result against the
    complex128 one (and against the exact basis state)."""
    import math

    from code_mojo_amd import _lib, lower_circuit
    from code_mojo_amd import workloads as W
    from code_mojo_amd.state import GpuState, Plan

    lib = _lib.load()
    n = args.qubits
    qc, v = W.value_encoding_precision(n, args.seed)
    ops, ud = lower_circuit(qc)
    rev = int(format(v, f"0{n}b")[::-1], 2)

    def timed(plan, g, steps, init):
        init(g)
        g.synchronize()
        for _ in range(max(args.warmup, 0)):
            plan.execute(g)
        g.synchronize()
        times = []
        for _ in range(steps):
            init(g)
            g.synchronize()
            t0 = time.perf_counter()
            plan.execute(g)
            g.synchronize()
            times.append(time.perf_counter() - t0)
        return min(times)

    reset = lambda g: g.reset()
    log(f"[bench] precision: {n}q value encoding, {len(ops)} gates")
    g64 = GpuState(n, np.float32)
    p_blk = Plan(n, np.float32, ops, ud, _lib.opts(mode=_lib.MODE_BLOCK, flags=_lib.FLAG_PROFILE))
    lib.bfly_profile_reset()
    t_blk = timed(p_blk, g64, max(args.steps, 1), reset)
    prof = profile_families(lib)
    blocks = p_blk.dump().count("\nblock ")
    g128 = GpuState(n, np.float64)
    p_fused = Plan(n, np.float64, ops, ud, _lib.opts())
    t_fused = timed(p_fused, g128, max(args.steps, 1), reset)
    up = GpuState(n, np.float64)
    up.copy_from(g64)
    l2 = math.sqrt(up.diff_norm2(g128))
    fid = abs(up.inner(g128)) ** 2
    fid_exact_64 = g64.prob_mask((1 << n) - 1, rev)
    fid_exact_128 = g128.prob_mask((1 << n) - 1, rev)
    g64.free()
    up.free()
    log(f"[bench] block {t_blk * 1e3:.2f} ms, fused c128 {t_fused * 1e3:.2f} ms, L2 {l2:.3g}, fidelity {fid:.12f}")
    # complex128 BLOCK mode (FP64 MFMA) on the same circuit
    p_blk128 = Plan(n, np.float64, ops, ud, _lib.opts(mode=_lib.MODE_BLOCK, flags=_lib.FLAG_PROFILE))
    lib.bfly_profile_reset()
    t_blk128 = timed(p_blk128, g128, max(args.steps, 1), reset)
    prof128 = profile_families(lib)
    fid_exact_blk128 = g128.prob_mask((1 << n) - 1, rev)
    log(f"[bench] block c128 {t_blk128 * 1e3:.2f} ms, fidelity vs exact {fid_exact_blk128:.15f}")
    p_plain = Plan(n, np.float64, ops, ud, _lib.opts(mode=_lib.MODE_UNFUSED))
    t_plain = timed(p_plain, g128, 1, reset)
    log(f"[bench] plain c128 {t_plain * 1e3:.1f} ms")
    g128.free()
    blk_ms, blk_n = prof.get("block_mfma_kernel", (0.0, 0))
    per_launch = blk_ms / blk_n if blk_n else float("nan")
    bytes_per = 2 * (1 << n) * 8
    flops_per = 256 * (1 << n)
    out = {
        "metric": "complex64 BLOCK-mode speedup over the plain complex128 GPU path at fidelity >= 1 - 1e-6",
        "value": round(t_plain / t_blk, 3), "unit": "x", "n_gpus": 1, "steps": max(args.steps, 1),
        "warmup": args.warmup, "ms_per_step": round(t_blk * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (MFMA) + f64 (phases, norm)",
        "data": "synthetic (seeded value-encoding integer)",
        "config": {"workload": "value_encoding_iqft_bitrev", "qubits": n, "value_seed": args.seed, "gates": len(ops),
                   "blocks": blocks, "block_qubits": 5, "renorm_every": 8},
        "block_c64_ms": round(t_blk * 1e3, 3), "fused_c128_ms": round(t_fused * 1e3, 3),
        "plain_c128_ms": round(t_plain * 1e3, 3), "speedup_vs_fused_c128": round(t_fused / t_blk, 3),
        "l2_vs_c128": l2, "fidelity_vs_c128": fid, "fidelity_vs_exact_c64": fid_exact_64,
        "fidelity_vs_exact_c128": fid_exact_128,
        "block_c128_ms": round(t_blk128 * 1e3, 3), "block_c128_fidelity_vs_exact": fid_exact_blk128,
        "block_c128_speedup_vs_fused_c128": round(t_fused / t_blk128, 3),
        "block_c128_kernel_mean_launch_ms": round(prof128.get("block_mfma_kernel", (0.0, 1))[0] /
                                                  max(prof128.get("block_mfma_kernel", (0.0, 1))[1], 1), 4),
        "roofline": {"bound": "hbm", "achieved": round(bytes_per / (per_launch * 1e-3) / 1e9, 2) if blk_n else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_per / (per_launch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if blk_n else None,
                     "traffic": read_traffic(f"block_ws_kernel_{n}q_value_encoding_traffic.txt"),
                     "kernel": "block_ws_kernel (complex64, 2^12 tiles)" if n >= 12 else "block_mfma_kernel",
                     "launches": blk_n,
                     "mean_launch_ms": round(per_launch, 4),
                     "mfma_tflops": round(flops_per / (per_launch * 1e-3) / 1e12, 2) if blk_n else None,
                     "mfma_peak_tflops": 157.3},
        "cpu_baseline": None,
    }
    return out


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary under profiles/ (None when there is none for this workload)."""
    tfile = ROOT / "profiles" / name
    if not tfile.exists():
        return None
    for line in tfile.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=", 1)[1])
    return None


def read_value(name: str, key: str):
    """`key = value` from a committed profile summary under profiles/ (None
    when the file or the key is missing)."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        line = line.strip().strip("`")
        if line.startswith(key) and "=" in line:
            try:
                return float(line.split("=", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def profile_families(lib) -> dict:
    """(total ms, launches) per kernel family from the library's event profile."""
    import ctypes as ct
    out = {}
    for idx in range(32):
        name, launches, ms = ct.c_char_p(), ct.c_int64(), ct.c_double()
        if lib.bfly_profile_get(idx, ct.byref(name), ct.byref(launches), ct.byref(ms)) != 0:
            break
        out[name.value.decode()] = (ms.value, launches.value)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--dtype", choices=["c128", "c64"], default="c128")
    ap.add_argument("--mode", choices=["fused", "unfused"], default="fused")
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--low-bits", type=int, default=0,
                    help="layered: low qubits kept in every tile (contiguous runs of 2^L amplitudes; 0 = default)")
    ap.add_argument("--max-gates", type=int, default=0, help="cap on gates per tile pass (0 = planner default)")
    ap.add_argument("--threads", type=int, default=0,
                    help="tile-pass workgroup size (0 = default; 2^(k-4) selects 16-amplitude register sweeps)")
    ap.add_argument("--jit", choices=["on", "off"], default="on",
                    help="layered: compile every tile pass into its own kernel at plan time (BFLY_FLAG_JIT; "
                         "the compile is part of plan building, outside the timed region)")
    ap.add_argument("--first-fit", action="store_true",
                    help="layered: tile passes take their qubits in program order (BFLY_FLAG_FIRST_FIT) "
                         "instead of the planner's search for the set that carries the most gates")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-qubits", type=int, default=0, help="CPU baseline qubits (0 = the headline's)")
    ap.add_argument("--cpu-layers", type=int, default=1, help="CPU baseline: the first L layers of the circuit")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="layered: skip the precision / batching / serving summaries of the default run")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batching: compile every circuit with the default options instead of the tuned ones")
    ap.add_argument("--workload", choices=["layered", "batching", "serving", "precision", "unitary"], default="layered",
                    help="layered: the headline state-GB/s workload (default); batching: many independent "
                         "circuits in batched persistent launches; serving: a Poisson request stream "
                         "through the serving engine")
    ap.add_argument("--circuits", type=int, default=0, help="batching / serving: request count "
                    "(default 4096 / 200)")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving: Poisson arrival rate (requests/s; 0 = 70 %% of the measured capacity)")
    ap.add_argument("--big", action="store_true", help="serving: add one 33-qubit complex128 QFT")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="N",
                    help="only copy one N-qubit complex128 state device to device (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes) and exit")
    ap.add_argument("--scale-probe", type=int, default=0, metavar="N",
                    help="only time the library's streaming kernel (bfly_scale: one in-place read and write of "
                         "every amplitude) on an N-qubit complex128 state, the measured HBM ceiling, and exit")
    args = ap.parse_args()
    if args.scale_probe:
        from code_mojo_amd.state import GpuState
        g = GpuState(args.scale_probe)
        g.fill_random(1)
        for _ in range(3):
            g.scale(1.0)
        g.synchronize()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            g.scale(1.0)
        g.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        nbytes = 2 * 2 * (1 << args.scale_probe) * 8
        print(json.dumps({"scale_probe_qubits": args.scale_probe, "bytes_per_launch": nbytes,
                          "ms_per_launch_wall": round(ms, 4), "read_write_gbs": round(nbytes / ms / 1e6, 1)}),
              flush=True)
        return 0
    if args.copy_probe:
        from code_mojo_amd.state import GpuState
        a, b = GpuState(args.copy_probe), GpuState(args.copy_probe)
        b.copy_from(a)
        b.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 2 * (1 << args.copy_probe) * 8,
                          "copy_probe_bytes_written": 2 * (1 << args.copy_probe) * 8,
                          "dispatches": 2}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the statevector strategy runs on a single device "
              f"(no sharding, no collectives); --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    from code_mojo_amd import _lib, lower_circuit
    from code_mojo_amd import workloads as W
    from code_mojo_amd.state import GpuState, Plan

    lib = _lib.load()
    if lib.bfly_device_count() < 1:
        log("no GPU visible")
        return 1
    if args.workload == "batching":
        print(json.dumps(bench_batching(args)), flush=True)
        return 0
    if args.workload == "serving":
        print(json.dumps(bench_serving(args)), flush=True)
        return 0
    if args.workload == "precision":
        print(json.dumps(bench_precision(args)), flush=True)
        return 0
    if args.workload == "unitary":
        print(json.dumps(bench_unitary(args)), flush=True)
        return 0
    out = bench_layered(args, lib)
    if not args.no_extras and args.qubits == 30 and args.dtype == "c128":
        out["extras"] = extras(args)
    print(json.dumps(out), flush=True)
    return 0


def bench_unitary(args, widths=(4, 5, 6, 7)) -> dict:
    """Dense m-qubit unitaries (UnitaryTransformation) on an n-qubit state:
    one Haar-random 2^m x 2^m block on the middle qubits per launch.  m = 4..6
    run on MFMA (unitary_mfma_kernel), wider ones on the FMA kernel
    (unitary_kernel).  Per launch: HBM rate of the in-place read + write of
    the state and the complex multiply-add rate (8 flops per complex MAC, 2^m
    MACs per amplitude), from the library's HIP-event profile."""
    from code_mojo_amd import _lib
    from code_mojo_amd.lowering import OpList
    from code_mojo_amd.state import GpuState, Plan

    lib = _lib.load()
    n = args.qubits if args.qubits <= 28 else 28
    rng = np.random.default_rng(args.seed)
    rows = {}
    for dname, dtype, b in (("c128", np.float64, 16), ("c64", np.float32, 8)):
        g = GpuState(n, dtype)
        g.fill_random(args.seed)
        for m in widths:
            z = rng.normal(size=(1 << m, 1 << m)) + 1j * rng.normal(size=(1 << m, 1 << m))
            q, r = np.linalg.qr(z)
            out = OpList(n)
            for _ in range(4):
                out.unitary(q * (np.diag(r) / np.abs(np.diag(r))), n // 2 - m // 2, m, [])
            ops, ud = out.finish()
            plan = Plan(n, dtype, ops, ud, _lib.opts(flags=_lib.FLAG_PROFILE))
            plan.execute(g)
            g.synchronize()
            lib.bfly_profile_reset()
            for _ in range(max(args.steps, 1)):
                plan.execute(g)
            g.synchronize()
            fam = profile_families(lib).get("unitary_kernel", (0.0, 0))
            plan.free()
            ms = fam[0] / max(fam[1], 1)
            gbs = 2 * (1 << n) * b / (ms * 1e-3) / 1e9
            tflops = (1 << n) * (1 << m) * 8 / (ms * 1e-3) / 1e12
            kernel = "unitary_mfma_kernel" if 4 <= m <= 6 else "unitary_kernel"
            rows[f"{dname}_m{m}"] = {"kernel": kernel, "ms_per_launch": round(ms, 4), "hbm_gbs": round(gbs, 1),
                                     "frac_of_hbm_peak": round(gbs / 8000.0, 4), "tflops": round(tflops, 2),
                                     "launches": fam[1]}
            log(f"[bench] unitary {dname} m={m}: {ms:.3f} ms per launch, {gbs:.0f} GB/s, {tflops:.1f} TFLOP/s "
                f"({kernel})")
        g.free()
    return {"metric": f"dense m-qubit unitary on a {n}-qubit state, per launch", "qubits": n, "results": rows,
            "peaks": {"hbm_gbs": 8000.0, "fp64_tflops": 78.6, "fp32_matrix_tflops": 157.3}}


def bench_host_boundary(args) -> dict:
    """The headline circuit through bfly_execute_host, the one-shot form a
    Mojo ExecutorFn binds: host re / im in (pageable numpy arrays), the state
    allocated, uploaded over PCIe, run and downloaded, in one call.  The
    PCIe-inclusive rate; never the headline value."""
    from code_mojo_amd import _lib, lower_circuit
    from code_mojo_amd import workloads as W
    from code_mojo_amd.state import Plan

    n = args.qubits
    ops, ud = lower_circuit(W.random_layered(n, args.layers, args.seed))
    opts = _lib.opts(flags=_lib.FLAG_JIT if args.jit == "on" else 0)
    passes = Plan(n, np.float64, ops, ud, opts).stats()["passes"]   # compiles (or reuses) the kernels
    re = np.zeros(1 << n)
    im = np.zeros(1 << n)
    re[:] = 0.0                       # touch every page before the timed call
    im[:] = 0.0
    re[0] = 1.0
    lib = _lib.load()
    t0 = time.perf_counter()
    _lib.check(lib.bfly_execute_host(n, 0, re.ctypes.data, im.ctypes.data, re.size, _lib.ops_ptr(ops), ops.size,
                                     _lib.udata_ptr(ud), ud.size // 2, opts))
    dt = time.perf_counter() - t0
    probe = float(np.sum(re[:1 << 20] ** 2 + im[:1 << 20] ** 2))
    moved = passes * 2 * (1 << n) * 16
    log(f"[bench] host boundary: {n}q execute_host {dt:.3f} s ({moved / dt / 1e9:.1f} GB/s state, PCIe included)")
    return {"what": f"{n}-qubit complex128 layered circuit via bfly_execute_host (allocate, upload, run, download, free)",
            "seconds": round(dt, 4), "state_gbps_pcie_inclusive": round(moved / dt / 1e9, 1),
            "state_bytes_each_way": 2 * (1 << n) * 16 // 2, "probe_prob_first_2^20": probe}


def extras(args) -> dict:
    """The other BASELINE configurations, measured in the same default run
    after the headline's timed region: the complex64 precision path (30-qubit
    value encoding), batching (4096 circuits) and serving (the seeded stream
    with the 33-qubit complex128 QFT, at 70 % of the measured capacity)."""
    import copy
    t0 = time.perf_counter()
    res = {}
    a = copy.copy(args)
    a.qubits, a.steps, a.warmup = 30, 1, 1
    p = bench_precision(a)
    res["precision"] = {k: p[k] for k in ("metric", "value", "unit", "block_c64_ms", "fused_c128_ms", "plain_c128_ms",
                                          "l2_vs_c128", "fidelity_vs_c128", "fidelity_vs_exact_c64",
                                          "block_c128_ms", "block_c128_fidelity_vs_exact")}
    res["precision"]["roofline"] = p["roofline"]
    log(f"[bench] extras: precision done at {time.perf_counter() - t0:.1f}s")
    b = copy.copy(args)
    b.circuits, b.steps = 4096, 1
    r = bench_batching(b)
    res["batching"] = {k: r[k] for k in ("metric", "value", "unit", "one_circuit_per_launch_circuits_per_s",
                                         "speedup_vs_one_circuit_per_launch", "min_fidelity_vs_c128",
                                         "checked_vs_c128", "max_amp_err_batch_vs_alone", "roofline")}
    log(f"[bench] extras: batching done at {time.perf_counter() - t0:.1f}s")
    c = copy.copy(args)
    c.circuits, c.rate, c.big = 200, 0.0, True
    v = bench_serving(c)
    res["serving"] = {k: v[k] for k in ("metric", "value", "unit", "one_circuit_per_launch_circuits_per_s",
                                        "speedup_vs_one_circuit_per_launch", "saturated", "at_rate",
                                        "small_requests_12_20q")}
    res["serving"]["config"] = v["config"]
    u = copy.copy(args)
    u.qubits, u.steps = 28, 3
    res["unitary"] = bench_unitary(u, widths=(4, 6))
    log(f"[bench] extras: unitary done at {time.perf_counter() - t0:.1f}s")
    res["host_boundary"] = bench_host_boundary(args)
    log(f"[bench] extras: host boundary done at {time.perf_counter() - t0:.1f}s")
    res["wall_s"] = round(time.perf_counter() - t0, 1)
    log(f"[bench] extras: serving done at {res['wall_s']}s")
    return res


def bench_layered(args, lib) -> dict:
    from code_mojo_amd import _lib, lower_circuit
    from code_mojo_amd import workloads as W
    from code_mojo_amd.state import GpuState, Plan

    n = args.qubits
    dtype = np.float64 if args.dtype == "c128" else np.float32
    b = 16 if args.dtype == "c128" else 8
    log(f"[bench] building {n}q x {args.layers} layer random layered circuit")
    t0 = time.perf_counter()
    qc = W.random_layered(n, args.layers, args.seed)
    ops, ud = lower_circuit(qc)
    mode = _lib.MODE_FUSED if args.mode == "fused" else _lib.MODE_UNFUSED
    flags = _lib.FLAG_PROFILE | (_lib.FLAG_JIT if args.jit == "on" else 0)
    flags |= _lib.FLAG_FIRST_FIT if args.first_fit else 0
    plan = Plan(n, dtype, ops, ud, _lib.opts(mode=mode, tile_bits=args.tile_bits, low_bits=args.low_bits,
                                         max_gates_per_pass=args.max_gates, threads=args.threads, flags=flags))
    plan_s = time.perf_counter() - t0
    st = plan.stats()
    log(f"[bench] lowered {len(ops)} gates -> {st['passes']} passes ({st['launches']} launches) "
        f"in {plan_s:.2f}s (jit {args.jit})")
    g = GpuState(n, dtype)
    log(f"[bench] state allocated ({g.nbytes() / 2**30:.1f} GiB); warmup {args.warmup}")
    for _ in range(args.warmup):
        plan.execute(g)
    g.synchronize()
    lib.bfly_profile_reset()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.execute(g)
        g.synchronize()
        log(f"[bench] step {i + 1}/{args.steps} done at {time.perf_counter() - t0:.3f}s")
    g.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3

    fam = {}
    for idx in range(16):
        name, launches, ms = _lib.ct.c_char_p(), _lib.ct.c_int64(), _lib.ct.c_double()
        if lib.bfly_profile_get(idx, _lib.ct.byref(name), _lib.ct.byref(launches), _lib.ct.byref(ms)) != 0:
            break
        if launches.value:
            fam[name.value.decode()] = (launches.value, ms.value)
    norm = g.norm2()
    log(f"[bench] kernel families: {fam}; norm^2 after run = {norm:.15f}")
    state_bytes = 2 * (1 << n) * b
    gbs = st["passes"] * state_bytes / (ms_per_step / 1e3) / 1e9
    dom = "tile_pass_kernel" if "tile_pass_kernel" in fam else max(fam, key=lambda k: fam[k][1])
    launches, dom_ms = fam[dom]
    per_launch_ms = dom_ms / launches
    achieved = state_bytes / (per_launch_ms / 1e3) / 1e9

    # the profile summaries are named for the kernel the profiler reports:
    # bfly_tile_pass for the generated tile-pass kernels, tile_pass_kernel for
    # the interpreting one
    kname = "bfly_tile_pass" if (args.jit == "on" and dom == "tile_pass_kernel") else dom
    traffic = read_traffic(f"{kname}_{n}q_layered_{args.dtype}_traffic.txt")
    fp64 = read_value(f"{kname}_{n}q_layered_{args.dtype}_pmc.md", "fp64_flops_per_launch")
    valu = read_value(f"{kname}_{n}q_layered_{args.dtype}_pmc.md", "valu_insts_per_launch")
    # VALU issue time of one launch: 4 cycles per wave64 instruction on one of
    # the 1024 SIMDs (256 CUs x 4) at the 2.4 GHz peak engine clock
    valu_issue_ms = valu * 4 / (1024 * 2.4e9) * 1e3 if valu else None
    # the same at the engine clock measured under this load (GRBM counters in
    # the pmc summary): the engine runs below its peak clock here
    clock_ghz = read_value(f"{kname}_{n}q_layered_{args.dtype}_pmc.md", "engine_clock_ghz_under_load")
    valu_issue_ms_clk = valu * 4 / (1024 * clock_ghz * 1e9) * 1e3 if (valu and clock_ghz) else None
    # the in-place read + write streaming ceiling measured on MI355X (a copy
    # kernel reading and writing 2 x 8 GiB in place), beside the 8 TB/s spec
    copy_gbs = read_value("hbm_copy_ceiling.md", "copy_read_write_gbs")

    out = {
        "metric": "state GB/s vs HBM peak at 30q",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.dtype == "c128" else "f32",
        "data": "synthetic (seeded random layered circuit, |0..0> start)",
        "config": {"workload": "random_layered_u3_czbrick_cx", "qubits": n, "layers": args.layers,
                   "state": "complex128" if args.dtype == "c128" else "complex64", "seed": args.seed,
                   "mode": args.mode, "tile_bits": args.tile_bits, "low_bits": args.low_bits, "max_gates_per_pass": args.max_gates,
                   "threads": args.threads, "tile_qubit_choice": "first fit" if args.first_fit else "search",
                   "gates": int(len(ops)), "passes_per_circuit": st["passes"],
                   "launches_per_circuit": st["launches"], "jit": args.jit == "on",
                   "plan_build_s": round(plan_s, 2)},
        "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
        "circuits_per_s": round(1e3 / ms_per_step, 4),
        "gate_layers_per_s": round(args.layers * 1e3 / ms_per_step, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname, "launches": launches, "mean_launch_ms": round(per_launch_ms, 4),
                     "bytes_per_launch": state_bytes,
                     "frac_of_measured_copy_ceiling": round(achieved / copy_gbs, 4) if copy_gbs else None,
                     "fp64_tflops": round(fp64 / (per_launch_ms * 1e-3) / 1e12, 2) if fp64 else None,
                     "fp64_frac_of_vector_peak": round(fp64 / (per_launch_ms * 1e-3) / 78.6e12, 4) if fp64 else None,
                     "valu_issue_frac": round(valu_issue_ms / per_launch_ms, 4) if valu_issue_ms else None,
                     "engine_clock_ghz_under_load": clock_ghz,
                     "valu_issue_frac_at_measured_clock": (round(valu_issue_ms_clk / per_launch_ms, 4)
                                                           if valu_issue_ms_clk else None)},
        "cpu_baseline": None,
    }
    if args.mode == "fused" and not args.first_fit and not args.no_extras:
        # The same circuit scheduled with first-fit tile qubits (program
        # order), timed on the same state after the timed region: the
        # searched schedule moves fewer bytes in fewer, heavier passes, so
        # state GB/s falls while the circuit runs faster.
        ff = Plan(n, dtype, ops, ud, _lib.opts(mode=mode, tile_bits=args.tile_bits, low_bits=args.low_bits,
                                               max_gates_per_pass=args.max_gates, threads=args.threads,
                                               flags=flags | _lib.FLAG_FIRST_FIT))
        ff.execute(g)
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ff.execute(g)
        g.synchronize()
        ff_ms = (time.perf_counter() - t0) / args.steps * 1e3
        ff_passes = ff.stats()["passes"]
        out["first_fit_schedule"] = {
            "what": "same circuit, tile qubits taken in program order (BFLY_FLAG_FIRST_FIT)",
            "passes_per_circuit": ff_passes, "ms_per_circuit": round(ff_ms, 3),
            "state_gbs": round(ff_passes * state_bytes / (ff_ms / 1e3) / 1e9, 2),
            "gate_layers_per_s": round(args.layers * 1e3 / ff_ms, 3),
            "searched_speedup": round(ff_ms / ms_per_step, 3)}
        log(f"[bench] first-fit schedule: {ff_passes} passes, {ff_ms:.1f} ms per circuit")
        ff.free()
    plan.free()
    g.free()
    if not args.no_cpu_baseline:
        # every core this process may use (the GPU box grants its share through OMP_NUM_THREADS)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0)) or 1
        cq = args.cpu_qubits or n
        log(f"[bench] cpu baseline: {cq}q x {args.cpu_layers} layers on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(cq, args.cpu_layers, args.seed, threads)
        log("[bench] cpu baseline done")
    return out


if __name__ == "__main__":
    sys.exit(main())
