"""GPU parity of circuits rewritten by the reference's fusion passes
(fuse_circuit / fuse_circuit_specialized_pairs, transformations_fusion.mojo:
879-908): FusedPairTransformation, UnitaryTransformation(m = 1, 2) and
ControlledUnitaryTransformation(m = 1) through the GPU strategy, against the
oracle run of the unfused circuit (complex128 max |error| <= 1e-12), in fused
mode (interpreted and generated tile kernels), unfused mode and BLOCK mode."""
import numpy as np
import pytest

from code_mojo_amd import _lib, algos, fusion, lower_circuit
from code_mojo_amd import workloads as W
from code_mojo_amd.state import GpuState, Plan, QuantumState
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_C128 = 1e-12


def circuit(name, n):
    if name == "qft":
        return W.qft_circuit(n), W.normal_state(n, 12)
    if name == "value":
        return algos.encode_value_circuit(n, 1234.0), QuantumState(n)
    return W.random_layered(n, 8, 31), QuantumState(n)


@pytest.mark.parametrize("name", ["qft", "value", "layered"])
@pytest.mark.parametrize("pass_", [fusion.fuse_circuit, fusion.fuse_circuit_specialized_pairs])
@pytest.mark.parametrize("mode,flags", [(_lib.MODE_FUSED, 0), (_lib.MODE_FUSED, _lib.FLAG_JIT),
                                        (_lib.MODE_UNFUSED, 0), (_lib.MODE_BLOCK, 0)])
def test_fused_circuits_vs_oracle(gpu_lib, name, pass_, mode, flags):
    n = 16
    qc, st = circuit(name, n)
    ref_ops, ref_ud = lower_circuit(qc)
    re, im = st.re.copy(), st.im.copy()
    O.execute(n, re, im, ref_ops, ref_ud)
    ops, ud = lower_circuit(pass_(qc))
    g = GpuState.from_host(st)
    p = Plan(n, np.float64, ops, ud, _lib.opts(mode=mode, flags=flags))
    p.execute(g)
    out = g.download()
    err = float(np.abs((out.re - re) + 1j * (out.im - im)).max())
    assert err <= TOL_C128, (name, pass_.__name__, mode, flags, err)
