"""GPU: circuits submitted as OpenQASM 3 text, lowered by the library's C++
front end, served by the engine and executed as plans: value encoding lands
on the exact basis state |v> (complex128 <= 1e-12), and a program with
measurements matches the oracle."""
import numpy as np
import pytest

from code_mojo_amd import _lib, algos, serve
from code_mojo_amd.state import GpuState, Plan
from oracle import oracle as O

from .test_qasm3 import qasm_from_circuit

pytestmark = pytest.mark.gpu


def test_value_encoding_from_qasm_through_the_engine(gpu_lib):
    jobs, values = [], []
    for i, n in enumerate((12, 16, 20)):
        v = 12345 % (1 << n) + i
        jobs.append(serve.qasm_job(f"ve{n}", qasm_from_circuit(algos.encode_value_circuit(n, float(v))), np.float64))
        values.append(v)
    with serve.Engine() as eng:
        res = serve.run_stream(jobs, None, eng, keep_results=True)
    assert res["completed"] == len(jobs)
    for job, v, st in zip(jobs, values, res["results"]):
        amp = st.re + 1j * st.im
        assert abs(abs(amp[v]) - 1) < 1e-12 and np.abs(np.delete(amp, v)).max() < 1e-12


def test_qasm_program_with_measurements_matches_oracle(gpu_lib):
    text = """OPENQASM 3.0;
include "stdgates.inc";
qubit[10] q;
qubit[4] anc;
bit[4] c;
h q[0]; h q[3]; h anc[2];
u3(0.3, 1.1, -0.4) q[7];
ctrl(2) @ rz(pi/3) q[0], q[3], q[9];
cry(2*pi/5) anc[2], q[5];
mcp(-pi/7) q[1], q[2], anc[0], q[8];
swap q[1], anc[3];
c[0] = measure q[3];
h q[3]; cx q[3], anc[1];
c = measure anc;
"""
    n, ops, ud = _lib.qasm3_lower(text)
    ops["gate"][ops["kind"] == 6] = 1            # seeded measurements for the comparison
    ops["controls"][ops["kind"] == 6] = 987654321
    re = np.zeros(1 << n)
    re[0] = 1
    im = np.zeros(1 << n)
    O.execute(n, re, im, ops, ud)
    g = GpuState(n)
    p = Plan(n, np.float64, ops, ud, _lib.opts())
    p.execute(g)
    got = g.download()
    assert max(np.abs(got.re - re).max(), np.abs(got.im - im).max()) <= 1e-12
    p.free()
    g.free()
