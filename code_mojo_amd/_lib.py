"""ctypes binding of libbfly_hip.so (the C ABI in include/bfly_hip.h).

The library is required: there is no CPU fallback.  Loading fails loudly if
the shared object is missing, and every non-zero status raises BflyError with
the library's thread-local message.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np

from .lowering import OP_DTYPE

LIB_PATH = Path(__file__).resolve().parent / "libbfly_hip.so"

BFLY_OK, BFLY_EINVAL, BFLY_ENOMEM, BFLY_EHIP, BFLY_EUNSUPPORTED, BFLY_ENUMERIC = 0, -1, -2, -3, -4, -5
C128, C64 = 0, 1
MODE_FUSED, MODE_UNFUSED, MODE_BLOCK, MODE_AUTO = 0, 1, 2, 3
TUNE_ALL = (1 << MODE_FUSED) | (1 << MODE_BLOCK) | (1 << MODE_AUTO)
FLAG_SYNC, FLAG_PROFILE, FLAG_GENERIC_PAIRS, FLAG_JIT, FLAG_FIRST_FIT = 1, 2, 4, 8, 16
BATCH_GRAPH = 1

ERROR_NAMES = {BFLY_EINVAL: "EINVAL", BFLY_ENOMEM: "ENOMEM", BFLY_EHIP: "EHIP",
               BFLY_EUNSUPPORTED: "EUNSUPPORTED", BFLY_ENUMERIC: "ENUMERIC"}


class BflyError(RuntimeError):
    """A libbfly_hip call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class ExecOpts(ct.Structure):
    _fields_ = [("mode", ct.c_int32), ("tile_bits", ct.c_int32), ("low_bits", ct.c_int32),
                ("block_qubits", ct.c_int32), ("max_gates_per_pass", ct.c_int32),
                ("renorm_every", ct.c_int32), ("threads", ct.c_int32), ("flags", ct.c_int32)]


class EngineOpts(ct.Structure):
    _fields_ = [("budget_bytes", ct.c_uint64), ("max_inflight_rounds", ct.c_int32), ("large_qubits", ct.c_int32),
                ("max_round_requests", ct.c_int32), ("flags", ct.c_int32), ("planner_threads", ct.c_int32),
                ("reserved", ct.c_int32)]


class RequestDesc(ct.Structure):
    _fields_ = [("num_qubits", ct.c_int32), ("dtype", ct.c_int32), ("ops", ct.c_void_p), ("nops", ct.c_uint64),
                ("udata", ct.c_void_p), ("udata_len", ct.c_uint64), ("init_re", ct.c_void_p),
                ("init_im", ct.c_void_p), ("init_kind", ct.c_int32), ("pad0", ct.c_int32),
                ("init_seed", ct.c_uint64), ("exec", ExecOpts)]


INIT_ZERO, INIT_HOST, INIT_RANDOM = 0, 1, 2
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)
BFLY_ETIMEDOUT, BFLY_ESTATE = -6, -7
ERROR_NAMES.update({BFLY_ETIMEDOUT: "ETIMEDOUT", BFLY_ESTATE: "ESTATE"})

_SIGS = {
    "bfly_last_error": (ct.c_char_p, []),
    "bfly_version": (ct.c_int, []),
    "bfly_device_count": (ct.c_int, []),
    "bfly_state_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int32, ct.c_int32]),
    "bfly_state_destroy": (ct.c_int, [ct.c_void_p]),
    "bfly_state_reset": (ct.c_int, [ct.c_void_p]),
    "bfly_state_upload": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint64]),
    "bfly_state_download": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint64]),
    "bfly_state_copy": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "bfly_state_info": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32)]),
    "bfly_state_device_ptrs": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p)]),
    "bfly_state_synchronize": (ct.c_int, [ct.c_void_p]),
    "bfly_plan_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_uint64,
                                    ct.c_void_p, ct.c_uint64, ct.POINTER(ExecOpts)]),
    "bfly_plan_destroy": (ct.c_int, [ct.c_void_p]),
    "bfly_plan_stats": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64),
                                   ct.POINTER(ct.c_int64)]),
    "bfly_plan_dump": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.c_uint64, ct.POINTER(ct.c_uint64)]),
    "bfly_plan_jit_info": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]),
    "bfly_plan_execute": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "bfly_apply": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_uint64, ct.c_void_p, ct.c_uint64,
                              ct.POINTER(ExecOpts)]),
    "bfly_execute_host": (ct.c_int, [ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_uint64, ct.c_void_p,
                                     ct.c_uint64, ct.c_void_p, ct.c_uint64, ct.POINTER(ExecOpts)]),
    "bfly_batch_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p),
                                     ct.c_int32, ct.c_int32]),
    "bfly_batch_launch": (ct.c_int, [ct.c_void_p]),
    "bfly_batch_synchronize": (ct.c_int, [ct.c_void_p]),
    "bfly_batch_stats": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]),
    "bfly_batch_destroy": (ct.c_int, [ct.c_void_p]),
    "bfly_engine_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.POINTER(EngineOpts)]),
    "bfly_engine_submit": (ct.c_int, [ct.c_void_p, ct.POINTER(RequestDesc), ct.POINTER(ct.c_uint64)]),
    "bfly_engine_status": (ct.c_int, [ct.c_void_p, ct.c_uint64, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double),
                                      ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "bfly_engine_wait": (ct.c_int, [ct.c_void_p, ct.c_uint64, ct.c_double]),
    "bfly_engine_fetch": (ct.c_int, [ct.c_void_p, ct.c_uint64, ct.c_void_p, ct.c_void_p, ct.c_uint64]),
    "bfly_engine_fetch_range": (ct.c_int, [ct.c_void_p, ct.c_uint64, ct.c_uint64, ct.c_uint64, ct.c_void_p,
                                            ct.c_void_p]),
    "bfly_engine_release": (ct.c_int, [ct.c_void_p, ct.c_uint64]),
    "bfly_engine_cancel": (ct.c_int, [ct.c_void_p, ct.c_uint64]),
    "bfly_engine_stats": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64),
                                     ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64), ct.POINTER(ct.c_uint64),
                                     ct.POINTER(ct.c_uint64)]),
    "bfly_engine_error": (ct.c_int, [ct.c_void_p, ct.c_uint64, ct.c_char_p, ct.c_uint64]),
    "bfly_engine_destroy": (ct.c_int, [ct.c_void_p]),
    "bfly_state_fill_random": (ct.c_int, [ct.c_void_p, ct.c_uint64]),
    "bfly_bitrev": (ct.c_int, [ct.c_void_p, ct.c_uint64]),
    "bfly_swap_qubits": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_int32]),
    "bfly_norm2": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_double)]),
    "bfly_prob_mask": (ct.c_int, [ct.c_void_p, ct.c_uint64, ct.c_uint64, ct.POINTER(ct.c_double)]),
    "bfly_prob_mask_noalloc": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_uint64, ct.c_uint64,
                                           ct.c_void_p, ct.c_void_p]),
    "bfly_marginal": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32), ct.c_int32, ct.POINTER(ct.c_double)]),
    "bfly_inner": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "bfly_diff_norm2": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_double)]),
    "bfly_scale": (ct.c_int, [ct.c_void_p, ct.c_double]),
    "bfly_qasm3_lower": (ct.c_int, [ct.c_char_p, ct.POINTER(ct.c_int32), ct.c_void_p, ct.c_uint64,
                                    ct.POINTER(ct.c_uint64), ct.c_void_p, ct.c_uint64, ct.POINTER(ct.c_uint64)]),
    "bfly_measure": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32), ct.c_int32, ct.c_int64,
                                ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_double)]),
    "bfly_sample": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32), ct.c_int32, ct.c_int64, ct.c_uint64,
                               ct.POINTER(ct.c_uint64)]),
    "bfly_autotune_candidates": (ct.c_int, [ct.c_int32, ct.c_int32, ct.POINTER(ExecOpts), ct.c_int32,
                                            ct.POINTER(ExecOpts), ct.c_int32, ct.POINTER(ct.c_int32)]),
    "bfly_autotune": (ct.c_int, [ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_uint64, ct.c_void_p, ct.c_uint64,
                                 ct.POINTER(ExecOpts), ct.c_int32, ct.c_int32, ct.POINTER(ExecOpts),
                                 ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "bfly_profile_reset": (ct.c_int, []),
    "bfly_profile_get": (ct.c_int, [ct.c_int32, ct.POINTER(ct.c_char_p), ct.POINTER(ct.c_int64),
                                    ct.POINTER(ct.c_double)]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load(path: os.PathLike | str | None = None) -> ct.CDLL:
    """Load (once) and return the library; raises if it was never built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP extension first (python -m code_mojo_amd.build); "
            "the GPU strategy has no CPU fallback")
    lib = ct.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != BFLY_OK:
        msg = load().bfly_last_error()
        raise BflyError(rc, msg.decode() if msg else "")


def opts(mode: int = MODE_FUSED, tile_bits: int = 0, low_bits: int = 0, block_qubits: int = 0,
         max_gates_per_pass: int = 0, renorm_every: int = 0, threads: int = 0, flags: int = 0) -> ExecOpts:
    return ExecOpts(mode, tile_bits, low_bits, block_qubits, max_gates_per_pass, renorm_every, threads, flags)


def ops_ptr(ops: np.ndarray) -> ct.c_void_p:
    assert ops.dtype == OP_DTYPE and ops.flags["C_CONTIGUOUS"]
    return ct.c_void_p(ops.ctypes.data) if ops.size else ct.c_void_p(0)


def udata_ptr(udata: np.ndarray) -> ct.c_void_p:
    assert udata.dtype == np.float64 and udata.flags["C_CONTIGUOUS"]
    return ct.c_void_p(udata.ctypes.data) if udata.size else ct.c_void_p(0)


def qasm3_lower(text: str):
    """(num_qubits, ops, udata) from OpenQASM 3 text, parsed and lowered by the
    library's C++ front end (bfly_qasm3_lower)."""
    lib = load()
    raw = text.encode("utf-8")
    n, nops, nud = ct.c_int32(), ct.c_uint64(), ct.c_uint64()
    check(lib.bfly_qasm3_lower(raw, ct.byref(n), None, 0, ct.byref(nops), None, 0, ct.byref(nud)))
    ops = np.zeros(nops.value, dtype=OP_DTYPE)
    udata = np.zeros(2 * nud.value, dtype=np.float64)
    check(lib.bfly_qasm3_lower(raw, ct.byref(n), ops.ctypes.data if ops.size else None, ops.size, ct.byref(nops),
                               udata.ctypes.data if udata.size else None, nud.value, ct.byref(nud)))
    return n.value, ops, udata
