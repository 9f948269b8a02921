"""Strategy registry and execute() (butterfly/core/executors.mojo:611-730).

`execute(state, circuit, ctx)` looks the strategy up in the registry and runs
it, as in the reference.  This build registers the GPU strategy
(ExecutionStrategy.GPU); the reference's CPU strategies are Mojo code that is
not part of this package, so selecting them raises like any unknown
strategy does in the reference (executors.mojo:715-720).

The GPU strategy lowers the whole transformation list once and crosses the
C ABI once per circuit.  A host QuantumState is uploaded, run and downloaded
in place (what the Mojo ExecutorFn shim does through bfly_execute_host); a
GpuState stays resident in HBM.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Union

from . import _lib
from .circuit import Circuit
from .context import GPU_MODES, ExecContext, ExecutionStrategy
from .lowering import lower_circuit
from .state import GpuState, Plan, QuantumState

ExecutorFn = Callable[[Union[QuantumState, GpuState], Circuit, ExecContext], None]


def exec_opts_from_ctx(ctx: ExecContext, flags: int = 0) -> _lib.ExecOpts:
    return _lib.opts(mode=GPU_MODES[ctx.gpu_mode], tile_bits=ctx.gpu_tile_bits, low_bits=ctx.gpu_low_bits,
                     block_qubits=ctx.gpu_fuse_k, max_gates_per_pass=ctx.gpu_max_gates_per_pass,
                     renorm_every=ctx.gpu_renorm_every, flags=flags)


def compile_plan(circuit: Circuit, dtype, ctx: ExecContext, flags: int = 0) -> Plan:
    ops, udata = lower_circuit(circuit, measure_seed=_measure_seed(ctx))
    return Plan(circuit.num_qubits, dtype, ops, udata, exec_opts_from_ctx(ctx, flags))


def _measure_seed(ctx: ExecContext):
    return ctx.measure_seed if ctx.measure_seed >= 0 else None


def execute_gpu(state: Union[QuantumState, GpuState], circuit: Circuit, ctx: ExecContext) -> None:
    """The GPU ExecutorFn."""
    if isinstance(state, GpuState):
        plan = compile_plan(circuit, state.dtype, ctx)
        if circuit.num_qubits != state.n:
            raise ValueError(f"Circuit qubits mismatch: circuit has {circuit.num_qubits}, state has {state.n}")
        plan.execute(state)
        state.synchronize()
        plan.free()
        return
    if circuit.num_qubits != state.n:
        raise ValueError(f"Circuit qubits mismatch: circuit has {circuit.num_qubits}, state has {state.n}")
    ops, udata = lower_circuit(circuit, measure_seed=_measure_seed(ctx))
    lib = _lib.load()
    opts = exec_opts_from_ctx(ctx)
    from .state import dtype_code
    _lib.check(lib.bfly_execute_host(state.n, dtype_code(state.dtype), state.re.ctypes.data, state.im.ctypes.data,
                                     state.size(), _lib.ops_ptr(ops), ops.size, _lib.udata_ptr(udata),
                                     udata.size // 2, opts))


@dataclass
class ExecutorEntry:
    strategy: int
    run: ExecutorFn


def executor_registry() -> List[ExecutorEntry]:
    return [ExecutorEntry(ExecutionStrategy.GPU, execute_gpu)]


def get_executor(strategy: int) -> ExecutorFn:
    for e in executor_registry():
        if e.strategy == strategy:
            return e.run
    raise ValueError(f"Unknown execution strategy: {strategy}")


def execute(state: Union[QuantumState, GpuState], circuit: Circuit, ctx: ExecContext | None = None) -> None:
    """executors.mojo:723-730."""
    c = (ctx or ExecContext()).copy()
    get_executor(c.execution_strategy)(state, circuit, c)
