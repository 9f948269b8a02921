"""The timed CPU baseline (oracle/grid_fused.c, the GRID_PARALLEL_FUSED
restatement) against the scalar oracle: the same amplitudes within rounding
on random gate / swap / bit-reversal lists and on the layered generator, for
the reference's default slack (3: eight rows), narrower and wider grids, and
one or several OpenMP threads.  Row-local groups, 256-column tiles of
cross-row gates, flat controlled passes and the three-CX swap all run."""
import numpy as np
import pytest

from code_mojo_amd import lower_circuit
from code_mojo_amd import workloads as W
from oracle import oracle as O

from .helpers import random_ops

TOL = 1e-13


def both(n, ops, ud, st, slack, threads):
    a_re, a_im = st.re.copy(), st.im.copy()
    O.execute(n, a_re, a_im, ops, ud)
    b_re, b_im = st.re.copy(), st.im.copy()
    O.grid_fused_execute(n, b_re, b_im, ops, slack=slack, threads=threads)
    return a_re + 1j * a_im, b_re + 1j * b_im


@pytest.mark.parametrize("n", [4, 8, 12, 16])
@pytest.mark.parametrize("slack,threads", [(3, 1), (3, 4), (1, 2), (5, 4), (-1, 0)])
def test_random_ops_match_the_scalar_oracle(n, slack, threads):
    rng = np.random.default_rng(500 + n)
    ops, ud = random_ops(rng, n, 90, with_blocks=False)
    st = W.normal_state(n, 3 + n)
    a, b = both(n, ops, ud, st, slack, threads)
    assert np.abs(a - b).max() <= TOL, (n, slack, threads)


@pytest.mark.parametrize("slack", [3, 4])
def test_layered_generator_matches(slack):
    n = 15
    ops, ud = lower_circuit(W.random_layered(n, 6, 2026))
    st = W.normal_state(n, 8)
    a, b = both(n, ops, ud, st, slack, 4)
    assert np.abs(a - b).max() <= TOL


def test_planted_wrong_qubit_is_caught():
    # the same comparison fails when one gate of the list acts on another qubit
    n = 10
    ops, ud = lower_circuit(W.random_layered(n, 2, 4))
    st = W.normal_state(n, 1)
    a, _ = both(n, ops, ud, st, 3, 2)
    bad = ops.copy()
    bad[3]["target"] = (bad[3]["target"] + 1) % n
    b_re, b_im = st.re.copy(), st.im.copy()
    O.grid_fused_execute(n, b_re, b_im, bad, slack=3, threads=2)
    assert np.abs(a - (b_re + 1j * b_im)).max() > 1e-3


def test_unsupported_ops_are_refused():
    n = 6
    rng = np.random.default_rng(1)
    ops, ud = random_ops(rng, n, 40, with_blocks=True, with_perms=False)
    if not (ops["kind"] == 3).any():
        pytest.skip("no dense block drawn")
    st = W.normal_state(n, 2)
    with pytest.raises(O.OracleError):
        O.grid_fused_execute(n, st.re.copy(), st.im.copy(), ops, slack=3, threads=1)
