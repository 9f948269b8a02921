"""The C ABI from a C program: tests/c/bfly_c_caller.c is compiled against
include/bfly_hip.h alone (no Python, no torch) and linked to libbfly_hip.so.

CPU: the program compiles warning-free as C11 against the header (the header's
structs keep their sizes: 104-byte bfly_op, 32-byte bfly_exec_opts, asserted
at compile time in the program).

GPU: the prebuilt program runs a 12-qubit circuit lowered by hand in C through
bfly_execute_host (complex128 and complex64), compiled plans on resident
states (generated kernels, unfused) and bfly_apply on an uploaded host state,
checks the error paths' codes itself, and writes its op array and results; the
oracle replays the same op records from the same start states.  complex128:
max |amplitude error| <= 1e-12; complex64: fidelity >= 1 - 1e-6 against the
complex128 oracle."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from code_mojo_amd.lowering import OP_DTYPE
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
CDIR = ROOT / "tests" / "c"
EXE = CDIR / "bfly_c_caller"


def test_c_caller_compiles_against_the_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    res = subprocess.run([cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-D_DEFAULT_SOURCE", "-fsyntax-only",
                          f"-I{ROOT / 'include'}", str(CDIR / "bfly_c_caller.c")], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def read_results(path: Path):
    raw = path.read_bytes()
    n, nops, nud, nruns = np.frombuffer(raw[:16], dtype=np.int32)
    off = 16
    ops = np.frombuffer(raw[off:off + nops * OP_DTYPE.itemsize], dtype=OP_DTYPE).copy()
    off += nops * OP_DTYPE.itemsize
    ud = np.frombuffer(raw[off:off + nud * 16], dtype=np.float64).copy()
    off += nud * 16
    size = 1 << int(n)
    rest = np.frombuffer(raw[off:], dtype=np.float64)
    assert rest.size == 2 * size * (nruns + 1)
    start = rest[:size] + 1j * rest[size:2 * size]
    runs = [rest[(2 + 2 * r) * size:(3 + 2 * r) * size] + 1j * rest[(3 + 2 * r) * size:(4 + 2 * r) * size]
            for r in range(nruns)]
    return int(n), ops, ud, start, runs


@pytest.mark.gpu
def test_c_caller_against_the_oracle(gpu_lib, tmp_path):
    assert EXE.exists(), "tests/c/bfly_c_caller is built by __graft_entry__.build()"
    out = tmp_path / "c_caller.bin"
    print(f"[c-caller] running {EXE.name}", flush=True)
    res = subprocess.run([str(EXE), str(out)], capture_output=True, text=True, timeout=120)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stderr
    assert "error paths ok" in res.stdout
    n, ops, ud, start, runs = read_results(out)
    kinds = set(int(k) for k in ops["kind"])
    assert {0, 1, 2, 3, 4, 5} <= kinds           # gates, bitrev, swap, unitary, phase flip, diffuser

    def oracle_from(st):
        re, im = st.real.copy(), st.imag.copy()
        O.execute(n, re, im, ops, ud)
        return re + 1j * im

    zero = np.zeros(1 << n, dtype=np.complex128)
    zero[0] = 1.0
    ref0 = oracle_from(zero)
    for r in (0, 1, 2):
        err = float(np.abs(runs[r] - ref0).max())
        assert err <= 1e-12, (r, err)
    ref4 = oracle_from(start)
    assert float(np.abs(runs[4] - ref4).max()) <= 1e-12
    a = runs[3]
    fid = abs(np.vdot(ref0, a)) ** 2 / (np.vdot(a, a).real * np.vdot(ref0, ref0).real)
    assert fid >= 1 - 1e-6, fid
