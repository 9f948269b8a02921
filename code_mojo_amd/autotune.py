"""Autotuner: tile shape and fusion width per circuit class.

The reference's strategies use fixed layouts (``compute_grid_layout``,
core/executors.mojo:213-228, 8 rows by default); the GPU strategy instead
times a few candidate schedules of a circuit on the device and keeps the
fastest.  The candidates and the timing are the library's
(``bfly_autotune_candidates`` / ``bfly_autotune``): FUSED tile passes at the
default tile size and one size either side, each without a cap and with a
cap of 2k gates per pass, the dense-block schedule (BLOCK) and the planner's
cost-model choice (AUTO), every one timed on a scratch state of the
circuit's size.  Results are cached per (family, qubits, dtype), the family
being the leading letters of the job name ("layered", "grover", "qft", ...),
so a batch or a request stream tunes each class once.
"""
from __future__ import annotations

import ctypes as ct
import re
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from ._lib import check
from .state import dtype_code

MODE_NAMES = {_lib.MODE_FUSED: "fused", _lib.MODE_UNFUSED: "unfused", _lib.MODE_BLOCK: "block",
              _lib.MODE_AUTO: "auto"}


def describe(o: _lib.ExecOpts) -> Dict[str, int | str]:
    """The options a tuned schedule differs in, as plain values."""
    return {"mode": MODE_NAMES.get(o.mode, str(o.mode)), "tile_bits": o.tile_bits,
            "max_gates_per_pass": o.max_gates_per_pass}


def candidates(n: int, dtype, base: Optional[_lib.ExecOpts] = None, modes: int = _lib.TUNE_ALL) -> List[_lib.ExecOpts]:
    """The schedules `tune` times for an n-qubit circuit (host only); the base
    options come first."""
    lib = _lib.load()
    b = base if base is not None else _lib.opts(mode=_lib.MODE_AUTO)
    count = ct.c_int32()
    check(lib.bfly_autotune_candidates(n, dtype_code(dtype), ct.byref(b), modes, None, 0, ct.byref(count)))
    out = (_lib.ExecOpts * max(count.value, 1))()
    check(lib.bfly_autotune_candidates(n, dtype_code(dtype), ct.byref(b), modes, out, count.value, ct.byref(count)))
    return [out[i] for i in range(count.value)]


@dataclass
class TuneResult:
    opts: _lib.ExecOpts
    best_ms: float
    base_ms: float
    seconds: float   # wall time spent tuning

    def as_dict(self) -> dict:
        return {**describe(self.opts), "ms": round(self.best_ms, 4), "base_ms": round(self.base_ms, 4)}


def tune(n: int, dtype, ops: np.ndarray, udata: np.ndarray, base: Optional[_lib.ExecOpts] = None,
         modes: int = _lib.TUNE_ALL, repeats: int = 3) -> TuneResult:
    """Time every candidate schedule of (ops, udata) on the GPU and return the
    fastest (median of `repeats` runs after one warm-up run each)."""
    lib = _lib.load()
    b = base if base is not None else _lib.opts(mode=_lib.MODE_AUTO)
    best = _lib.ExecOpts()
    best_ms, base_ms = ct.c_double(), ct.c_double()
    t0 = time.perf_counter()
    check(lib.bfly_autotune(n, dtype_code(dtype), _lib.ops_ptr(ops), len(ops), _lib.udata_ptr(udata), udata.size // 2,
                            ct.byref(b), modes, repeats, ct.byref(best), ct.byref(best_ms), ct.byref(base_ms)))
    return TuneResult(best, best_ms.value, base_ms.value, time.perf_counter() - t0)


def family(name: str) -> str:
    m = re.match(r"[A-Za-z_]+", name)
    return m.group(0) if m else name


class Autotuner:
    """Per-class cache of tuned options.

    `opts_for(job)` tunes the job's class (family, qubits, dtype) on first
    sight and returns the cached choice afterwards.  Circuits wider than
    `max_qubits` are not tuned (the scratch state would double their
    footprint): they get `base`."""

    def __init__(self, modes: int = _lib.TUNE_ALL, repeats: int = 3, max_qubits: int = 26,
                 base: Optional[_lib.ExecOpts] = None):
        self.modes = modes
        self.repeats = repeats
        self.max_qubits = max_qubits
        self.base = base if base is not None else _lib.opts(mode=_lib.MODE_AUTO)
        self.cache: Dict[Tuple[str, int, str], TuneResult] = {}

    def key(self, job) -> Tuple[str, int, str]:
        return family(job.name), job.n, np.dtype(job.dtype).name

    def opts_for(self, job) -> _lib.ExecOpts:
        if job.n > self.max_qubits:
            return self.base
        k = self.key(job)
        r = self.cache.get(k)
        if r is None:
            r = tune(job.n, job.dtype, job.ops, job.udata, self.base, self.modes, self.repeats)
            self.cache[k] = r
        return r.opts

    def seconds(self) -> float:
        return sum(r.seconds for r in self.cache.values())

    def summary(self) -> Dict[str, dict]:
        return {f"{f}/{n}/{d}": r.as_dict() for (f, n, d), r in sorted(self.cache.items())}
