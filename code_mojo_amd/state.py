"""Host and device states.

QuantumState mirrors the reference's State (butterfly/core/state.mojo:19-98):
two SoA arrays `re`, `im` of length 2^n, initialised to |0..0>.  The float
dtype plays the role of the reference's compile-time FloatType
(butterfly/core/types.mojo:7): float64 = complex128 path, float32 = the
`-D FLOAT32` complex64 path.

GpuState owns a device-resident copy (libbfly_hip state handle); it stays in
HBM between executions, which is what the serving loop relies on.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check


def dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return _lib.C128
    if dt == np.float32:
        return _lib.C64
    raise ValueError(f"unsupported amplitude dtype {dt}")


def np_dtype(code: int):
    return np.float64 if code == _lib.C128 else np.float32


class QuantumState:
    """Host SoA state (butterfly/core/state.mojo:19-98)."""

    def __init__(self, n: int, dtype=np.float64, re: Optional[np.ndarray] = None, im: Optional[np.ndarray] = None):
        self.dtype = np.dtype(dtype)
        dtype_code(self.dtype)
        if re is None:
            self.re = np.zeros(1 << n, dtype=self.dtype)
            self.im = np.zeros(1 << n, dtype=self.dtype)
            self.re[0] = 1.0
        else:
            self.re = np.ascontiguousarray(re, dtype=self.dtype)
            self.im = np.ascontiguousarray(im, dtype=self.dtype)
            if self.re.shape != (1 << n,) or self.im.shape != (1 << n,):
                raise ValueError("state arrays must have length 2^n")
        self.n = n

    @classmethod
    def from_complex(cls, amps: np.ndarray, dtype=np.float64) -> "QuantumState":
        amps = np.asarray(amps)
        n = int(amps.size).bit_length() - 1
        if (1 << n) != amps.size:
            raise ValueError("State size must be a power of two")
        return cls(n, dtype, amps.real.copy(), amps.imag.copy())

    def size(self) -> int:
        return self.re.size

    def __len__(self) -> int:
        return self.size()

    def __getitem__(self, idx: int) -> complex:
        return complex(self.re[idx], self.im[idx])

    def __setitem__(self, idx: int, val: complex) -> None:
        self.re[idx] = val.real
        self.im[idx] = val.imag

    def to_complex(self) -> np.ndarray:
        return self.re.astype(np.float64) + 1j * self.im.astype(np.float64)

    def copy(self) -> "QuantumState":
        return QuantumState(self.n, self.dtype, self.re.copy(), self.im.copy())


class GpuState:
    """Device-resident state (bfly_state handle)."""

    def __init__(self, n: int, dtype=np.float64):
        lib = _lib.load()
        self.n = n
        self.dtype = np.dtype(dtype)
        self.code = dtype_code(self.dtype)
        h = ct.c_void_p()
        check(lib.bfly_state_create(ct.byref(h), n, self.code))
        self._h = h

    @property
    def handle(self) -> ct.c_void_p:
        if self._h is None:
            raise RuntimeError("state was freed")
        return self._h

    @classmethod
    def from_host(cls, state: QuantumState) -> "GpuState":
        g = cls(state.n, state.dtype)
        g.upload(state)
        return g

    def free(self) -> None:
        if getattr(self, "_h", None) is not None:
            _lib.load().bfly_state_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def nbytes(self) -> int:
        return 2 * (1 << self.n) * self.dtype.itemsize

    def reset(self) -> None:
        check(_lib.load().bfly_state_reset(self.handle))

    def upload(self, state: QuantumState) -> None:
        if state.n != self.n or state.dtype != self.dtype:
            raise ValueError("host / device state shape mismatch")
        check(_lib.load().bfly_state_upload(self.handle, state.re.ctypes.data, state.im.ctypes.data,
                                            state.size()))

    def upload_arrays(self, re: np.ndarray, im: np.ndarray) -> None:
        self.upload(QuantumState(self.n, self.dtype, re, im))

    def download(self, out: Optional[QuantumState] = None) -> QuantumState:
        if out is None:
            out = QuantumState(self.n, self.dtype)
        check(_lib.load().bfly_state_download(self.handle, out.re.ctypes.data, out.im.ctypes.data, out.size()))
        return out

    def synchronize(self) -> None:
        check(_lib.load().bfly_state_synchronize(self.handle))

    def copy_from(self, other: "GpuState") -> None:
        check(_lib.load().bfly_state_copy(self.handle, other.handle))

    # -- permutations and reductions ----------------------------------------
    def bitrev(self, targets: Sequence[int] = ()) -> None:
        mask = 0
        for t in targets:
            mask |= 1 << int(t)
        check(_lib.load().bfly_bitrev(self.handle, mask))

    def swap_qubits(self, a: int, b: int) -> None:
        check(_lib.load().bfly_swap_qubits(self.handle, a, b))

    def norm2(self) -> float:
        out = ct.c_double()
        check(_lib.load().bfly_norm2(self.handle, ct.byref(out)))
        return out.value

    def prob_mask(self, mask: int, value: int) -> float:
        out = ct.c_double()
        check(_lib.load().bfly_prob_mask(self.handle, mask, value, ct.byref(out)))
        return out.value

    def marginal(self, targets: Sequence[int]) -> np.ndarray:
        m = len(targets)
        t = (ct.c_int32 * m)(*targets)
        out = np.zeros(1 << m, dtype=np.float64)
        check(_lib.load().bfly_marginal(self.handle, t, m, out.ctypes.data_as(ct.POINTER(ct.c_double))))
        return out

    def inner(self, other: "GpuState") -> complex:
        """<self|other> = sum conj(self) * other."""
        r, i = ct.c_double(), ct.c_double()
        check(_lib.load().bfly_inner(self.handle, other.handle, ct.byref(r), ct.byref(i)))
        return complex(r.value, i.value)

    def diff_norm2(self, other: "GpuState") -> float:
        out = ct.c_double()
        check(_lib.load().bfly_diff_norm2(self.handle, other.handle, ct.byref(out)))
        return out.value

    def scale(self, s: float) -> None:
        check(_lib.load().bfly_scale(self.handle, float(s)))

    def measure(self, targets: Sequence[int], values: Optional[Sequence[Optional[bool]]] = None,
                seed: Optional[int] = None) -> Tuple[int, float]:
        """apply_measure (state.mojo:447-553) on the device state: collapse onto
        an outcome of the free targets chosen with the reference's MINSTD draw
        from `seed` (None: the clock).  Returns (outcome, its probability);
        bit k of the outcome is the k-th free target."""
        m = len(targets)
        vals = list(values) if values else [None] * m
        if len(vals) != m:
            raise ValueError("Measurement values size mismatch")
        t = (ct.c_int32 * max(m, 1))(*targets)
        v = (ct.c_int32 * max(m, 1))(*[-1 if x is None else int(bool(x)) for x in vals])
        out, prob = ct.c_uint64(), ct.c_double()
        check(_lib.load().bfly_measure(self.handle, t, v, m, -1 if seed is None else int(seed),
                                       ct.byref(out), ct.byref(prob)))
        return int(out.value), float(prob.value)

    def sample(self, targets: Sequence[int], shots: int, seed: int) -> np.ndarray:
        """`shots` outcomes of `targets` without collapse (MINSTD sequence from
        `seed`; shot 0 is the outcome measure() would choose)."""
        m = len(targets)
        t = (ct.c_int32 * max(m, 1))(*targets)
        out = np.zeros(max(int(shots), 1), dtype=np.uint64)
        check(_lib.load().bfly_sample(self.handle, t, m, int(seed), int(shots),
                                      out.ctypes.data_as(ct.POINTER(ct.c_uint64))))
        return out[:int(shots)]

    def fill_random(self, seed: int) -> None:
        """re, im ~ N(0,1) from `seed` (SplitMix64 + Box-Muller), normalised."""
        check(_lib.load().bfly_state_fill_random(self.handle, int(seed)))


class Plan:
    """A compiled pass schedule for one circuit (bfly_plan)."""

    def __init__(self, n: int, dtype, ops: np.ndarray, udata: np.ndarray, exec_opts: _lib.ExecOpts):
        lib = _lib.load()
        self.n = n
        self.code = dtype_code(dtype)
        self._ops, self._udata = ops, udata
        h = ct.c_void_p()
        check(lib.bfly_plan_create(ct.byref(h), n, self.code, _lib.ops_ptr(ops), ops.size,
                                   _lib.udata_ptr(udata), udata.size // 2, ct.byref(exec_opts)))
        self._h = h

    def stats(self) -> dict:
        p, l, g = ct.c_int64(), ct.c_int64(), ct.c_int64()
        check(_lib.load().bfly_plan_stats(self._h, ct.byref(p), ct.byref(l), ct.byref(g)))
        return {"passes": p.value, "launches": l.value, "gates": g.value}

    def jit_info(self) -> dict:
        """Generated tile-pass kernels of a BFLY_FLAG_JIT plan and how many
        were compiled for it (the rest came from the process cache)."""
        k, c = ct.c_int64(), ct.c_int64()
        check(_lib.load().bfly_plan_jit_info(self._h, ct.byref(k), ct.byref(c)))
        return {"kernels": k.value, "compiled": c.value}

    def dump(self) -> str:
        """Text form of the compiled schedule (bfly_plan_dump)."""
        lib = _lib.load()
        need = ct.c_uint64()
        check(lib.bfly_plan_dump(self._h, None, 0, ct.byref(need)))
        buf = ct.create_string_buffer(need.value)
        check(lib.bfly_plan_dump(self._h, buf, need.value, ct.byref(need)))
        return buf.value.decode()

    def execute(self, state: GpuState) -> None:
        check(_lib.load().bfly_plan_execute(self._h, state.handle))

    def free(self) -> None:
        if getattr(self, "_h", None) is not None:
            _lib.load().bfly_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many circuits executed round by round (bfly_batch): round r runs step
    r of every plan, tile passes of one round grouped into one launch per
    tile size.  graph=True captures all rounds into one hipGraph."""

    def __init__(self, states, plans, graph: bool = False):
        if len(states) != len(plans) or not states:
            raise ValueError("one plan per state")
        lib = _lib.load()
        self._states, self._plans = list(states), list(plans)  # keep alive
        sa = (ct.c_void_p * len(states))(*[s.handle.value for s in states])
        pa = (ct.c_void_p * len(plans))(*[p._h.value for p in plans])
        h = ct.c_void_p()
        check(lib.bfly_batch_create(ct.byref(h), sa, pa, len(states), _lib.BATCH_GRAPH if graph else 0))
        self._h = h

    def launch(self) -> None:
        check(_lib.load().bfly_batch_launch(self._h))

    def synchronize(self) -> None:
        check(_lib.load().bfly_batch_synchronize(self._h))

    def stats(self) -> dict:
        r, l = ct.c_int64(), ct.c_int64()
        check(_lib.load().bfly_batch_stats(self._h, ct.byref(r), ct.byref(l)))
        return {"rounds": r.value, "launches": l.value}

    def free(self) -> None:
        if getattr(self, "_h", None) is not None:
            _lib.load().bfly_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
