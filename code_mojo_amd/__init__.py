"""code_mojo_amd -- MI355X (gfx950) GPU execution strategy for the Butterfly
statevector simulator (learnqc/code_mojo).

The package mirrors the reference's circuit / state / strategy API so that a
circuit built as in Butterfly runs on the GPU by selecting
ExecutionStrategy.GPU.  All amplitude work happens in libbfly_hip.so
(hand-written HIP kernels behind a C ABI, include/bfly_hip.h); this Python
layer only builds circuits, lowers them to the op array and calls the ABI.
"""
from .circuit import (Circuit, ClassicalReplacementKind, ClassicalTransformation, ControlKind, ControlledUnitaryTransformation,
                      FusedPairTransformation, GateTransformation, MeasurementTransformation,
                      QuantumCircuit, QuantumRegister, QubitReversalTransformation, Register,
                      SwapTransformation, UnitaryTransformation, bit_reverse, measure,
                      permute_qubits, replace_quantum_with_classical)
from .context import ExecContext, ExecutionStrategy, execution_strategy_name
from .executors import compile_plan, execute, execute_gpu, executor_registry, get_executor
from .fusion import fuse_circuit, fuse_circuit_specialized_pairs
from .gates import (Custom_Gate, GateInfo, GateKind, H, P, P_Gate, RX, RX_Gate, RY, RY_Gate, RZ,
                    RZ_Gate, U1, U2, U3, X, Y, Z, H_Gate, X_Gate, Y_Gate, Z_Gate)
from .lowering import LoweringError, OP_DTYPE, lower_circuit
from .qasm3 import QasmError, parse_qasm3_file, parse_qasm3_string
from .state import GpuState, Plan, QuantumState

__all__ = [name for name in dir() if not name.startswith("_")]
