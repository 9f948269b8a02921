"""CPU tests of BFLY_FLAG_JIT plans: every tile pass is generated as its own
kernel and compiled with hipRTC for gfx950 when the plan is created (no GPU
needed), and a second plan of the same structure takes every kernel from the
process cache."""
import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd.state import Plan

from .helpers import random_ops


def _plan(n, dtype, ops, ud, **kw):
    return Plan(n, dtype, ops, ud, _lib.opts(flags=_lib.FLAG_JIT, **kw))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_jit_plan_compiles_every_tile_pass(dtype):
    n = 14
    ops, ud = lower_circuit(W.random_layered(n, 4, 11))
    p = _plan(n, dtype, ops, ud, tile_bits=10)
    info = p.jit_info()
    assert info["kernels"] == p.stats()["passes"] > 1
    assert info["compiled"] <= info["kernels"]


def test_jit_cache_reuses_kernels_of_the_same_structure():
    n = 13
    ops, ud = lower_circuit(W.random_layered(n, 3, 5))
    first = _plan(n, np.float64, ops, ud, tile_bits=9).jit_info()
    second = _plan(n, np.float64, ops, ud, tile_bits=9).jit_info()
    assert second["kernels"] == first["kernels"] > 0
    assert second["compiled"] == 0
    # new angles, same structure: the generated sources do not depend on the
    # gate parameters, only on kinds, slots, forms and masks
    ops2, ud2 = lower_circuit(W.random_layered(n, 3, 6))
    assert _plan(n, np.float64, ops2, ud2, tile_bits=9).jit_info()["kernels"] > 0


@pytest.mark.parametrize("n,tile_bits,k", [(6, 4, 4), (9, 6, 6), (12, 0, 12), (13, 11, 11)])
def test_jit_handles_every_instruction_form(n, tile_bits, k):
    # random op lists: controlled pairs, diagonal gates on lane / tile bits,
    # phase flips, scale records and both sweep widths (threads 2^(k-4):
    # 16-amplitude sweeps; the default is 8-amplitude sweeps from k = 10)
    rng = np.random.default_rng(40 + n)
    ops, ud = random_ops(rng, n, 80)
    for threads in (0, 1 << (k - 4) if k >= 10 else 64):
        p = _plan(n, np.float64, ops, ud, tile_bits=tile_bits, threads=threads)
        assert p.jit_info()["kernels"] == sum(1 for line in p.dump().splitlines() if line.startswith("tile "))


def test_plans_without_the_flag_report_no_kernels():
    n = 10
    ops, ud = lower_circuit(W.qft_circuit(n))
    p = Plan(n, np.float64, ops, ud, _lib.opts())
    assert p.jit_info() == {"kernels": 0, "compiled": 0}
    pu = Plan(n, np.float64, ops, ud, _lib.opts(mode=_lib.MODE_UNFUSED, flags=_lib.FLAG_JIT))
    assert pu.jit_info()["kernels"] == 0


def _dump_sources(tmp_path, n, dtype, env, monkeypatch, **kw):
    monkeypatch.setenv("BFLY_JIT_DUMP", str(tmp_path))
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    ops, ud = lower_circuit(W.random_layered(n, 40, 2026))
    p = _plan(n, dtype, ops, ud, **kw)
    srcs = [(tmp_path / f"pass{i}.hip").read_text() for i in range(p.jit_info()["kernels"])]
    for f in tmp_path.glob("*.hip"):
        f.unlink()
    return srcs


def _direct_sides(src):
    # a staged pass loads the tile into pr / pi and stores it from LDS through
    # the stage-in offsets; a direct side reads or writes the first / last
    # sweep's members at base + gj
    return ("pr[0] = *" not in src,
            "*reinterpret_cast<V*>(im + base + toff" not in src)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_direct_stage_in_and_out_only_in_passes_of_six_sweeps(tmp_path, monkeypatch, dtype):
    n = 20
    srcs = _dump_sources(tmp_path, n, dtype, {}, monkeypatch)
    direct = 0
    for s in srcs:
        d_in, d_out = _direct_sides(s)
        sweeps = s.count("// sweep ")
        if d_in or d_out:
            assert sweeps >= 6
            direct += 1
        if d_in:
            assert "gj =" in s.split("// sweep 1:")[0]
    assert direct > 0
    off = _dump_sources(tmp_path, n, dtype, {"BFLY_JIT_DIRECT": "0"}, monkeypatch)
    assert len(off) == len(srcs) and not any(any(_direct_sides(s)) for s in off)
    every = _dump_sources(tmp_path, n, dtype, {"BFLY_JIT_DIRECT": "1", "BFLY_JIT_DIRECT_MIN_SWEEPS": "1"},
                          monkeypatch)
    assert all(all(_direct_sides(s)) for s in every)


def test_launch_forms_by_tile_count_and_width(tmp_path, monkeypatch):
    # 2^12-amplitude tiles: complex128 passes of 6 or more sweeps move re and
    # im through the LDS in separate round trips (a one-component tile) and
    # run one block per tile; the other passes keep 16-byte round trips, and
    # at 24 qubits (4096 tiles per pass) take the work-queue form (tile
    # counter argument, atomic take), at 23 qubits not; BFLY_JIT_GRID=tile
    # keeps one block per tile; complex64 passes of 4096 tiles take the queue
    monkeypatch.setenv("BFLY_JIT_DUMP", str(tmp_path))

    def sources(n, dtype=np.float64, **env):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ops, ud = lower_circuit(W.random_layered(n, 12, 7))
        p = _plan(n, dtype, ops, ud, tile_bits=12)
        out = [(tmp_path / f"pass{i}.hip").read_text() for i in range(p.jit_info()["kernels"])]
        for f in tmp_path.glob("*.hip"):
            f.unlink()
        for k in env:
            monkeypatch.delenv(k)
        return out

    d24 = sources(24)
    heavy = [s for s in d24 if s.count("// sweep ") >= 6]
    light = [s for s in d24 if s.count("// sweep ") < 6]
    assert heavy and all("T s_c[" in s and "atomicAdd" not in s for s in heavy)
    assert all("C2 s_c[" in s and "atomicAdd(Q, 1ull)" in s for s in light)
    q24 = sources(24, BFLY_JIT_SPLIT="0")
    assert q24 and all("atomicAdd(Q, 1ull)" in s and "unsigned long long* __restrict__ Q" in s and "C2 s_c[" in s
                       for s in q24)
    assert not any("atomicAdd" in s for s in sources(23, BFLY_JIT_SPLIT="0"))
    assert not any("atomicAdd" in s for s in sources(24, BFLY_JIT_SPLIT="0", BFLY_JIT_GRID="tile"))
    forced = sources(14, BFLY_JIT_SPLIT="0", BFLY_JIT_QUEUE_MIN_TILES="1")
    assert forced and all("atomicAdd(Q, 1ull)" in s for s in forced)
    c64 = sources(25, np.float32)
    assert c64 and all("atomicAdd(Q, 1ull)" in s for s in c64)


def test_split_round_trip_structure(tmp_path, monkeypatch):
    # complex128 passes of 6 or more sweeps hold one component in LDS: each
    # sweep boundary is re out, barrier, re in, barrier, the previous layout's
    # im out, barrier, im in (three barriers), and the launch bound asks for
    # three waves per SIMD; complex64 passes keep 16-byte elements
    srcs = _dump_sources(tmp_path, 20, np.float64, {}, monkeypatch)
    split = [s for s in srcs if "T s_c[" in s]
    assert split
    for s in split:
        nsw = s.count("// sweep ")
        assert nsw >= 6
        assert "__launch_bounds__(256, 3)" in s
        # one im write at the previous layout per boundary after the first sweep
        assert s.count("const unsigned jbp = padc(") >= nsw - 1
        assert s.count("__syncthreads()") >= 3 * (nsw - 1)
    for s in srcs:
        if "T s_c[" not in s:
            assert s.count("// sweep ") < 6 and "C2 s_c[" in s
    f32 = _dump_sources(tmp_path, 20, np.float32, {}, monkeypatch)
    assert f32 and all("C2 s_c[" in s for s in f32)
