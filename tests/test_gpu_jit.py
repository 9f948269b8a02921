"""GPU parity of BFLY_FLAG_JIT plans (tile passes compiled into their own
straight-line kernels): against the oracle at the drop-in's bars (complex128
max |amplitude error| <= 1e-12; complex64 against the complex128 oracle with
L2 <= 2e-5), and against the interpreting tile kernel (the same factored
arithmetic; the two differ in where the compilers contract products into FMAs
and in the generated kernels' member factors, products of a sweep's phases
rounded once in FP64: <= 1e-14 in complex128 at 18 qubits)."""
import math

import numpy as np
import pytest

from code_mojo_amd import _lib, algos, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd.state import GpuState, Plan, QuantumState
from oracle import oracle as O

from .helpers import random_ops

pytestmark = pytest.mark.gpu

TOL_C128 = 1e-12
TOL_JIT_VS_INTERP = 1e-14


def run(n, ops, ud, st, dtype=np.float64, jit=True, **kw):
    g = GpuState(n, dtype)
    g.upload(QuantumState(n, dtype, st.re.astype(dtype), st.im.astype(dtype)))
    p = Plan(n, dtype, ops, ud, _lib.opts(flags=_lib.FLAG_JIT if jit else 0, **kw))
    p.execute(g)
    out = g.download()
    info = p.jit_info()
    p.free()
    g.free()
    return out.re.astype(np.float64) + 1j * out.im.astype(np.float64), info


def oracle(n, ops, ud, st):
    re, im = st.re.astype(np.float64).copy(), st.im.astype(np.float64).copy()
    O.execute(n, re, im, ops, ud)
    return re + 1j * im


@pytest.mark.parametrize("n,tile_bits,threads", [(5, 4, 0), (9, 6, 0), (12, 8, 64), (14, 0, 0), (16, 12, 0),
                                                 (16, 12, 256), (17, 10, 0), (17, 10, 64), (18, 13, 0),
                                                 (21, 13, 0)])
def test_random_ops_vs_oracle(gpu_lib, n, tile_bits, threads):
    rng = np.random.default_rng(300 + n)
    ops, ud = random_ops(rng, n, 70)
    st = W.normal_state(n, 9 + n)
    out, info = run(n, ops, ud, st, tile_bits=tile_bits, threads=threads)
    assert info["kernels"] > 0
    err = np.abs(out - oracle(n, ops, ud, st)).max()
    assert err <= TOL_C128, (n, tile_bits, threads, err)


@pytest.mark.parametrize("name", ["layered", "qft", "value", "grover"])
def test_config_circuits_vs_oracle(gpu_lib, name):
    n = 18
    if name == "layered":
        qc, st = W.random_layered(n, 12, 3), QuantumState(n)
    elif name == "qft":
        qc, st = W.qft_circuit(n), W.normal_state(n, 4)
    elif name == "value":
        qc, st = algos.encode_value_circuit(n, 123457.0), QuantumState(n)
    else:
        n = 12
        qc, _ = W.grover_search(n, 77)
        st = QuantumState(n)
    ops, ud = lower_circuit(qc)
    out, info = run(n, ops, ud, st)
    assert info["kernels"] > 0
    ref = oracle(n, ops, ud, st)
    assert np.abs(out - ref).max() <= TOL_C128
    interp, _ = run(n, ops, ud, st, jit=False)
    assert np.abs(out - interp).max() <= TOL_JIT_VS_INTERP


@pytest.mark.parametrize("tile_bits", [0, 11])
def test_complex64_vs_complex128_oracle(gpu_lib, tile_bits):
    n = 17
    ops, ud = lower_circuit(W.random_layered(n, 10, 21))
    st = QuantumState(n)
    out, info = run(n, ops, ud, st, dtype=np.float32, tile_bits=tile_bits)
    assert info["kernels"] > 0
    ref = oracle(n, ops, ud, st)
    l2 = float(np.linalg.norm(out - ref))
    fid = float(abs(np.vdot(ref, out)) ** 2 / (np.vdot(out, out).real * np.vdot(ref, ref).real))
    assert l2 <= 2e-5 and fid >= 1 - 1e-6, (l2, fid)


def test_repeated_execution_and_cached_kernels(gpu_lib):
    n = 15
    ops, ud = lower_circuit(W.random_layered(n, 8, 8))
    st = W.normal_state(n, 2)
    a, first = run(n, ops, ud, st)
    b, second = run(n, ops, ud, st)
    assert second["compiled"] == 0 and second["kernels"] == first["kernels"]
    assert np.array_equal(a, b)   # the same kernels on the same input: identical bits
    # one plan executed twice equals the circuit applied twice
    g = GpuState(n)
    g.upload(st)
    p = Plan(n, np.float64, ops, ud, _lib.opts(flags=_lib.FLAG_JIT))
    p.execute(g)
    p.execute(g)
    twice = g.download()
    ref = oracle(n, ops, ud, st)
    ref2 = oracle(n, ops, ud, QuantumState(n, np.float64, ref.real.copy(), ref.imag.copy()))
    assert np.abs(twice.re + 1j * twice.im - ref2).max() <= TOL_C128


def test_26q_layered_jit_equals_interpreter(gpu_lib):
    # at scale: the full 40-layer generator, JIT against the interpreting
    # kernel (size-independent agreement) and the norm
    n = 26
    ops, ud = lower_circuit(W.random_layered(n, 40, 2026))
    st = QuantumState(n)
    a, info = run(n, ops, ud, st)
    assert info["kernels"] > 10
    b, _ = run(n, ops, ud, st, jit=False)
    assert np.abs(a - b).max() <= 1e-13
    assert abs(float(np.vdot(a, a).real) - 1.0) <= 1e-12
