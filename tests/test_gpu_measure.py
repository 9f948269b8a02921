"""GPU: measurement and sampling through the C ABI against the oracle.
complex128: collapsed state within 1e-12 of the oracle's; complex64 within
2e-6.  The outcome probabilities are FP64 compensated sums on the device and
the reference's naive sums in the oracle, so a draw landing within ~1e-15 of
an outcome boundary could pick a neighbour; the seeds here do not.  complex64
with more than 12 distinct free qubits is checked against the library's own
draw rule and the oracle's collapse (see test_measure_matches_oracle)."""
import numpy as np
import pytest

from code_mojo_amd import _lib, lower_circuit, measure, serve
from code_mojo_amd import workloads as W
from code_mojo_amd.lowering import OpList
from code_mojo_amd.state import Batch, GpuState, Plan
from oracle import oracle as O

from .plan_sim import apply_measure

pytestmark = pytest.mark.gpu
M = 2147483647


def oracle_measure(n, st, targets, values, seed, dtype=np.float64):
    out = OpList(n)
    out.measure(targets, [None if v < 0 else bool(v) for v in values], seed)
    ops, ud = out.finish()
    re, im = st.re.astype(dtype).copy(), st.im.astype(dtype).copy()
    O.execute(n, re, im, ops, ud)
    return re, im


def gpu_state(n, st, dtype):
    g = GpuState(n, dtype)
    g.upload_arrays(st.re.astype(dtype), st.im.astype(dtype))
    return g


CASES = [
    (10, [3], [-1], 1234567),
    (12, [0, 5, 11], [-1, -1, -1], 98765431),
    (14, [2, 9, 4, 13], [-1, 1, -1, 0], 55555),
    (16, [15, 0], [-1, -1], 2 ** 35 + 17),
    (17, list(range(6)), [-1] * 6, 424242424),
    (18, [1, 3, 5, 7, 9, 11, 13, 15, 17, 0, 2, 4], [-1] * 12, 31337),   # 12 free: 4096 outcomes
    (13, [4, 4, 7], [-1, -1, -1], 777777),                             # a repeated target
    (12, [6, 2, 6], [1, -1, -1], 13579135),                            # free and forced on one qubit
    # more than 12 distinct free qubits: one thread per outcome and rest chunk
    (14, list(range(14)), [-1] * 14, 8675309),                          # every qubit: no rest bits
    (16, [15, 1, 8, 3, 12, 0, 6, 10, 5, 13, 2, 9, 11, 7], [-1] * 13 + [0], 2468013579),
    (18, list(range(17, 1, -1)), [-1] * 16, 112358132),                # 16 free, reversed order
    (20, list(range(20)), [-1] * 20, 99991),                           # 2^20 outcomes
]


def fp64_outcome_probs(n, re, im, targets, values):
    """Outcome probabilities of the free targets in FP64 (bit k = k-th free target)."""
    idx = np.arange(1 << n, dtype=np.int64)
    p = re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2
    keep = np.ones(1 << n, dtype=bool)
    out = np.zeros(1 << n, dtype=np.int64)
    k = 0
    for t, v in zip(targets, values):
        bit = (idx >> t) & 1
        if v >= 0:
            keep &= bit == v
        else:
            out |= bit << k
            k += 1
    return np.bincount(out[keep], weights=p[keep], minlength=1 << k)


def distinct_free(targets, values):
    forced = {t for t, v in zip(targets, values) if v >= 0}
    return len({t for t, v in zip(targets, values) if v < 0 and t not in forced})


@pytest.mark.parametrize("n,targets,values,seed", CASES)
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_measure_matches_oracle(gpu_lib, n, targets, values, seed, dtype, tol):
    st = W.normal_state(n, n + seed % 97)
    g = gpu_state(n, st, dtype)
    outcome, prob = g.measure(targets, [None if v < 0 else bool(v) for v in values], seed)
    got = g.download()
    if dtype == np.float32 and distinct_free(targets, values) > 12:
        # Thousands of outcomes in float32: the reference's naive running sums
        # carry errors of the order of one outcome's probability, so which
        # outcome r lands in depends on the summation.  Check the library's
        # stated rule instead (FP64 outcome sums, running sums and r in
        # float32), the probability, and the collapse against the oracle's
        # collapse onto the chosen outcome.
        re32, im32 = st.re.astype(dtype), st.im.astype(dtype)
        P = fp64_outcome_probs(n, re32, im32, targets, values)
        cum = np.cumsum(P.astype(np.float32))
        rng = (seed % M) * 48271 % M
        r = np.float32(rng) / np.float32(M) * cum[-1]
        assert outcome == int(np.searchsorted(cum, r, side="left")), (outcome, prob)
        assert abs(prob - P[outcome]) <= 1e-6 * P[outcome], (prob, P[outcome])   # returned in float32
        fi = 0
        forced = []
        for v in values:
            if v < 0:
                forced.append((outcome >> fi) & 1)
                fi += 1
            else:
                forced.append(v)
        re_o, im_o = oracle_measure(n, st, targets, forced, seed, dtype)
    else:
        re_o, im_o = oracle_measure(n, st, targets, values, seed, dtype)
    assert max(np.abs(got.re - re_o).max(), np.abs(got.im - im_o).max()) <= tol, (outcome, prob)
    assert abs(g.norm2() - 1.0) < (1e-12 if dtype == np.float64 else 1e-5)
    g.free()


def test_reference_errors_on_gpu(gpu_lib):
    n = 6
    g = GpuState(n)                     # |0...0>
    with pytest.raises(_lib.BflyError, match="Measurement probability is zero"):
        g.measure([2], [True], 5)
    with pytest.raises(_lib.BflyError, match="Target out of bounds"):
        g.measure([6], None, 5)
    one = GpuState(1)
    one.upload_arrays(np.array([0.0, 1.0]), np.zeros(2))
    with pytest.raises(_lib.BflyError, match="outcome probability is zero"):
        one.measure([0], None, 0)       # seed 0 draws r = 0: outcome 0 with p = 0
    with pytest.raises(_lib.BflyError, match="more than 24 sampled targets"):
        GpuState(25).sample(list(range(25)), 10, 3)
    g.free()
    one.free()


@pytest.mark.parametrize("mode", [_lib.MODE_FUSED, _lib.MODE_UNFUSED])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_mid_circuit_measurement_in_plans(gpu_lib, mode, dtype, tol):
    n = 16
    qc = W.random_layered(n, 4, 5)
    measure(qc, [2, 7, 12], [None, None, False], seed=1234567)
    for q in range(0, n, 2):
        qc.h(q)
    measure(qc, [0, 1, 2, 3, 4], seed=7654321)
    ops, ud = lower_circuit(qc)
    st = W.normal_state(n, 4)
    re_o, im_o = st.re.copy(), st.im.copy()
    O.execute(n, re_o, im_o, ops, ud)
    g = gpu_state(n, st, dtype)
    p = Plan(n, dtype, ops, ud, _lib.opts(mode=mode))
    p.execute(g)
    got = g.download()
    assert max(np.abs(got.re - re_o).max(), np.abs(got.im - im_o).max()) <= tol
    p.free()
    g.free()


def test_sample_shot0_and_histogram(gpu_lib):
    n = 14
    st = W.normal_state(n, 21)
    g = gpu_state(n, st, np.float64)
    targets = [13, 0, 6]
    seed = 192837465
    shots = g.sample(targets, 40000, seed)
    assert np.array_equal(shots, g.sample(targets, 40000, seed))          # deterministic
    # shot 0 is the outcome apply_measure picks with the same seed
    psi = st.to_complex()
    collapsed = apply_measure(psi, n, targets, [-1, -1, -1], seed)
    idx = np.arange(1 << n)
    o = ((idx >> 13) & 1) | (((idx >> 0) & 1) << 1) | (((idx >> 6) & 1) << 2)
    assert set(np.unique(o[np.abs(collapsed) > 0])) == {int(shots[0])}
    # histogram against the marginal (binomial 5-sigma bars)
    p = np.array([np.sum(np.abs(psi[o == k]) ** 2) for k in range(8)])
    counts = np.bincount(shots.astype(np.int64), minlength=8)
    sigma = np.sqrt(40000 * p * (1 - p))
    assert np.all(np.abs(counts - 40000 * p) <= 5 * sigma + 1)
    out_before = g.download()
    assert np.array_equal(out_before.re, st.re)                            # sampling does not collapse
    with pytest.raises(_lib.BflyError):
        g.sample(targets, 10, M)                                          # r stuck at 0
    g.free()


def test_batches_and_engine_reject_measurement(gpu_lib):
    n = 10
    qc = W.random_layered(n, 2, 3)
    measure(qc, [1], seed=99)
    ops, ud = lower_circuit(qc)
    g = GpuState(n)
    p = Plan(n, np.float64, ops, ud, _lib.opts())
    with pytest.raises(_lib.BflyError, match="not in batches"):
        Batch([g], [p])
    job = serve.Job(name="m", n=n, dtype=np.float64, ops=ops, udata=ud)
    with serve.Engine() as eng:
        rid = eng.submit(job)
        eng.wait(rid, 60)
        assert eng.status(rid)[0] == _lib.REQ_FAILED
        assert "measurement" in eng.error(rid)
        eng.release(rid)
    p.free()
    g.free()


def test_sample_many_targets(gpu_lib):
    """16 sampled targets (2^16 outcomes, the per-outcome kernel): shot 0 is
    the outcome measure() picks with the same seed, and every shot lands on
    an outcome with nonzero probability of the FP64 marginal."""
    n = 18
    st = W.normal_state(n, 23)
    # zero the amplitudes with qubit 5 set: outcomes with that bit have p = 0
    idx = np.arange(1 << n)
    re, im = st.re.copy(), st.im.copy()
    re[(idx >> 5) & 1 == 1] = 0.0
    im[(idx >> 5) & 1 == 1] = 0.0
    nrm = np.sqrt(np.sum(re ** 2 + im ** 2))
    re, im = re / nrm, im / nrm
    g = GpuState(n)
    g.upload_arrays(re, im)
    targets = [17, 5, 0, 1, 2, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14]
    assert len(set(targets)) == 16
    seed = 5550123
    shots = g.sample(targets, 5000, seed)
    P = fp64_outcome_probs(n, re, im, targets, [-1] * 16)
    assert np.all(P[shots.astype(np.int64)] > 0)
    assert not np.any((shots >> np.uint64(1)) & np.uint64(1))       # qubit 5 (bit 1) never set
    outcome, _ = g.measure(targets, None, seed)
    assert outcome == int(shots[0])
    g.free()


WIDE = [
    (26, list(range(26)), [-1] * 26, 20261204),                          # every qubit: 2^26 outcomes
    # 25 distinct free qubits in falling order, one of them listed twice,
    # and one forced qubit: 26 free entries
    (26, list(range(25, 0, -1)) + [7, 0], [-1] * 26 + [1], 734019),
]


@pytest.mark.parametrize("n,targets,values,seed", WIDE)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_measure_more_than_24_free_targets(gpu_lib, n, targets, values, seed, dtype):
    """More than 24 free targets (the reference allocates 2^f outcome sums,
    state.mojo:475-501): the library searches the outcome level by level.
    The draw is made in FP64 for both precisions (a float32 running sum over
    2^25 or more outcomes stagnates: single outcome probabilities fall below
    half its ulp).  complex128: the collapsed state within 1e-12 of the
    oracle's (which keeps the 2^f sums); complex64: the outcome has the
    probability returned, lies where r = rng / M * total falls in the FP64
    running sums of the float32 state's outcome probabilities (within 1e-9 of
    the total), and the collapse equals the oracle's collapse onto it."""
    print(f"[measure-wide] n={n} free={sum(v < 0 for v in values)} seed={seed} {np.dtype(dtype).name}", flush=True)
    st = W.normal_state(n, seed % 1000)
    g = gpu_state(n, st, dtype)
    outcome, prob = g.measure(targets, [None if v < 0 else bool(v) for v in values], seed)
    got = g.download()
    assert abs(g.norm2() - 1.0) < (1e-12 if dtype == np.float64 else 1e-5)
    g.free()
    free = [t for t, v in zip(targets, values) if v < 0]
    forced = []
    fi = 0
    for t, v in zip(targets, values):
        if v < 0:
            forced.append((outcome >> fi) & 1)
            fi += 1
        else:
            forced.append(v)
    # the outcome is consistent: a repeated free qubit carries one value
    for k, t in enumerate(free):
        for k2, t2 in enumerate(free):
            if t == t2:
                assert ((outcome >> k) & 1) == ((outcome >> k2) & 1)
    if dtype == np.float64:
        re_o, im_o = oracle_measure(n, st, targets, values, seed)
    else:
        re32, im32 = st.re.astype(dtype), st.im.astype(dtype)
        P = fp64_outcome_probs(n, re32, im32, targets, values)
        assert P[outcome] > 0 and abs(prob - P[outcome]) <= 1e-6 * P[outcome], (prob, P[outcome])
        cum = np.cumsum(P)
        total = cum[-1]
        r = ((seed % M) * 48271 % M) / M * total
        lo = cum[outcome] - P[outcome]
        assert lo - 1e-9 * total <= r <= cum[outcome] + 1e-9 * total, (outcome, r, lo, cum[outcome])
        del P, cum
        re_o, im_o = oracle_measure(n, st, targets, forced, seed, dtype)
    tol = 1e-12 if dtype == np.float64 else 2e-6
    assert max(np.abs(got.re - re_o).max(), np.abs(got.im - im_o).max()) <= tol, (outcome, prob)
