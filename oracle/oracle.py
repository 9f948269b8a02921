"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker or the timed CPU baseline.  It runs the
CPU restatement of the reference's SCALAR strategy (bfly_oracle.c) and of
GRID_PARALLEL_FUSED (grid_fused.c) over the same lowered op array the GPU
strategy consumes.
"""
from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> Path:
    """Compile liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    res = subprocess.run(["make", "-C", str(HERE), "-B", "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB


def load() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ct.CDLL(str(LIB))
        for sfx, ftype in (("f64", ct.c_double), ("f32", ct.c_float)):
            fn = getattr(lib, f"oracle_execute_{sfx}")
            fn.restype = ct.c_int
            fn.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint64, ct.c_void_p,
                           ct.c_uint64, ct.c_char_p, ct.c_int]
            fn = getattr(lib, f"oracle_bitrev_{sfx}")
            fn.restype = ct.c_int
            fn.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_uint64]
        lib.grid_fused_execute_f64.restype = ct.c_int
        lib.grid_fused_execute_f64.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint64,
                                               ct.c_int, ct.c_int, ct.c_int]
        lib.grid_fused_max_threads.restype = ct.c_int
        _lib = lib
    return _lib


class OracleError(RuntimeError):
    pass


def execute(n: int, re: np.ndarray, im: np.ndarray, ops: np.ndarray, udata: np.ndarray) -> None:
    """Run the scalar-strategy restatement in place on (re, im)."""
    lib = load()
    sfx = "f64" if re.dtype == np.float64 else "f32"
    assert re.dtype == im.dtype and re.flags["C_CONTIGUOUS"] and im.flags["C_CONTIGUOUS"]
    assert re.size == 1 << n
    err = ct.create_string_buffer(256)
    rc = getattr(lib, f"oracle_execute_{sfx}")(
        n, re.ctypes.data, im.ctypes.data, ops.ctypes.data if ops.size else None, ops.size,
        udata.ctypes.data if udata.size else None, udata.size // 2, err, 256)
    if rc != 0:
        raise OracleError(f"oracle error {rc}: {err.value.decode()}")


def bitrev(n: int, re: np.ndarray, im: np.ndarray, mask: int = 0) -> None:
    lib = load()
    sfx = "f64" if re.dtype == np.float64 else "f32"
    rc = getattr(lib, f"oracle_bitrev_{sfx}")(n, re.ctypes.data, im.ctypes.data, mask)
    if rc != 0:
        raise OracleError(f"oracle bitrev error {rc}")


def grid_fused_execute(n: int, re: np.ndarray, im: np.ndarray, ops: np.ndarray, slack: int = -1,
                       min_bits: int = -1, threads: int = 0) -> None:
    """GRID_PARALLEL_FUSED restatement (OpenMP) in place, float64 only."""
    lib = load()
    assert re.dtype == np.float64
    rc = lib.grid_fused_execute_f64(n, re.ctypes.data, im.ctypes.data, ops.ctypes.data if ops.size else None,
                                    ops.size, slack, min_bits, threads)
    if rc != 0:
        raise OracleError(f"grid-fused error {rc}")


def max_threads() -> int:
    return load().grid_fused_max_threads()
