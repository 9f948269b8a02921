"""Autotuner candidate lists (host only, no GPU): the base options first, the
FUSED tile sizes around the default with and without a gates-per-pass cap,
BLOCK only from 9 qubits, AUTO; duplicates removed; tile sizes never wider
than the state unless they are the default; mode masks respected."""
import numpy as np
import pytest

from code_mojo_amd import _lib, autotune


def sig(o):
    return (o.mode, o.tile_bits, o.max_gates_per_pass)


def test_full_candidate_list_complex128():
    c = [sig(o) for o in autotune.candidates(24, np.float64)]
    assert c[0] == (_lib.MODE_AUTO, 0, 0)   # the base options come first
    fused = {(k, g) for m, k, g in c if m == _lib.MODE_FUSED}
    # default tile 2^12 is spelled 0
    assert fused == {(11, 0), (11, 22), (0, 0), (0, 24), (13, 0), (13, 26)}
    assert (_lib.MODE_BLOCK, 0, 0) in c
    assert len(c) == len(set(c)) == 1 + 6 + 1   # AUTO equals the base: kept once


def test_complex64_default_tile_and_small_states():
    c = [sig(o) for o in autotune.candidates(24, np.float32)]
    # default 2^13 (spelled 0) is the widest tile: 2^12 is the only other size
    assert {(k, g) for m, k, g in c if m == _lib.MODE_FUSED} == {(12, 0), (12, 24), (0, 0), (0, 26)}
    # a state no wider than the default tile also gets tiles narrower than itself
    # (a 2^12 tile on a 12-qubit state is the whole state: spelled 0)
    c12 = [sig(o) for o in autotune.candidates(12, np.float32)]
    assert {k for m, k, g in c12 if m == _lib.MODE_FUSED} == {0, 11, 10}
    small = [sig(o) for o in autotune.candidates(8, np.float64)]
    assert {k for m, k, g in small if m == _lib.MODE_FUSED} == {0, 7, 6}
    assert all(m != _lib.MODE_BLOCK for m, k, g in small)   # BLOCK needs 9 qubits
    assert all(k <= 8 or k == 0 for m, k, g in small)


@pytest.mark.parametrize("modes,allowed", [
    (1 << _lib.MODE_FUSED, {_lib.MODE_FUSED}),
    (1 << _lib.MODE_BLOCK, {_lib.MODE_BLOCK}),
    (0, set()),
])
def test_mode_mask(modes, allowed):
    base = _lib.opts(mode=_lib.MODE_FUSED, tile_bits=12)
    c = autotune.candidates(20, np.float64, base=base, modes=modes)
    assert sig(c[0]) == (_lib.MODE_FUSED, 0, 0)   # base first (its default tile spelled 0)
    assert {o.mode for o in c[1:]} <= allowed


def test_family_keys():
    assert autotune.family("layered20x20") == "layered"
    assert autotune.family("grover14") == "grover"
    assert autotune.family("qft30") == "qft"
