"""Writes the golden fixtures under tests/golden/ from closed forms (numpy).

The reference (Mojo) cannot be compiled or run anywhere here and holds no
data fixtures of its own, so the vectors come from mathematics that is
independent of any pair-walk kernel:

  dense_small.npz  random op lists over every gate kind, single / multi
                   controls, dense blocks, swaps and bit reversals on
                   n = 1..6; expected output = full 2^n x 2^n matrix product
                   built with Kronecker products of the gate definitions
                   (butterfly/core/gates.mojo:5-61 conventions: qubit q = bit
                   q, LSB first; control test (i & mask) == mask)
  qft_fft.npz      QFT of N(0,1) states for n = 2..12: sqrt(N) * ifft(psi)
  bitrev.npz       full and partial bit-reversal permutation tables from
                   integer bit manipulation (state.mojo:131-204,252-301)
  grover.npz       sin^2((2k+1) theta), theta = asin(2^-n/2) for n = 3..10
                   (algos/grover_classical.mojo:154-168)

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from code_mojo_amd import gates as G  # noqa: E402
from code_mojo_amd.lowering import OP_DTYPE, OpList  # noqa: E402
from code_mojo_amd.workloads import normal_state  # noqa: E402


def kron_all(mats):
    out = np.array([[1.0 + 0j]])
    for m in mats:
        out = np.kron(out, m)
    return out


P0 = np.array([[1, 0], [0, 0]], dtype=complex)
P1 = np.array([[0, 0], [0, 1]], dtype=complex)
I2 = np.eye(2, dtype=complex)


def dense_gate(n, target, controls, g):
    """(I - Pc) + Pc (g on target), Pc = projector on all controls = 1."""
    def op(with_gate):
        mats = []
        for q in range(n - 1, -1, -1):  # kron order: qubit n-1 leftmost
            if q == target:
                mats.append(g if with_gate else I2)
            elif q in controls:
                mats.append(P1)
            else:
                mats.append(I2)
        return kron_all(mats)
    dim = 1 << n
    return np.eye(dim, dtype=complex) - op(False) + op(True)


def dense_unitary(n, target, m, u, controls):
    mats = [np.eye(1 << (n - target - m))] if n - target - m > 0 else []
    blk = kron_all(mats + [u] + ([np.eye(1 << target)] if target > 0 else []))
    if not controls:
        return blk
    diag = np.ones(1 << n)
    idx = np.arange(1 << n)
    mask = sum(1 << c for c in controls)
    sel = (idx & mask) == mask
    # controlled: act with blk where controls are set (controls are outside the block)
    out = np.diag((~sel).astype(complex))
    out += np.diag(sel.astype(complex)) @ blk @ np.diag(sel.astype(complex))
    del diag
    return out


def perm_matrix(n, perm):
    """out[i] = in[perm[i]]."""
    dim = 1 << n
    M = np.zeros((dim, dim), dtype=complex)
    M[np.arange(dim), perm] = 1
    return M


def bitrev_perm(n, mask):
    idx = np.arange(1 << n)
    if mask == 0:
        mask = (1 << n) - 1
    pos = [b for b in range(n) if (mask >> b) & 1]
    k = len(pos)
    out = idx & ~mask
    for i, b in enumerate(pos):
        out |= ((idx >> b) & 1) << pos[k - 1 - i]
    return out


def swap_perm(n, a, b):
    idx = np.arange(1 << n)
    x = ((idx >> a) & 1) ^ ((idx >> b) & 1)
    return idx ^ ((x << a) | (x << b))


def random_unitary(rng, dim):
    z = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_case(rng, n, nops):
    ops = OpList(n)
    U = np.eye(1 << n, dtype=complex)
    kinds = ["H", "P", "X", "Y", "Z", "RX", "RY", "RZ", "U3"]
    for _ in range(nops):
        r = rng.random()
        if r < 0.70 or n < 2:
            t = int(rng.integers(n))
            others = [q for q in range(n) if q != t]
            nc = int(rng.integers(0, min(3, len(others)) + 1)) if others else 0
            ctrls = [int(c) for c in rng.choice(others, size=nc, replace=False)] if nc else []
            k = kinds[int(rng.integers(len(kinds)))]
            th = float(rng.uniform(-2 * math.pi, 2 * math.pi))
            if k == "U3":
                info = G.Custom_Gate(G.U3(th, float(rng.uniform(0, 6.3)), float(rng.uniform(0, 6.3))))
            else:
                kind = G.GateKind.names.index(k)
                info = G.GateInfo.of(kind, th if kind in (1, 5, 6, 7) else None)
            ops.gate(t, ctrls, info.kind, info.gate, info.arg or 0.0)
            U = dense_gate(n, t, ctrls, info.gate) @ U
        elif r < 0.82:
            m = int(rng.integers(1, min(3, n) + 1))
            t = int(rng.integers(0, n - m + 1))
            outside = [q for q in range(n) if not (t <= q < t + m)]
            nc = int(rng.integers(0, min(2, len(outside)) + 1)) if outside else 0
            ctrls = [int(c) for c in rng.choice(outside, size=nc, replace=False)] if nc else []
            u = random_unitary(rng, 1 << m)
            ops.unitary(u, t, m, ctrls)
            U = dense_unitary(n, t, m, u, ctrls) @ U
        elif r < 0.91:
            a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
            ops.swap(a, b)
            U = perm_matrix(n, swap_perm(n, a, b)) @ U
        else:
            if rng.random() < 0.5:
                mask = 0
            else:
                k = int(rng.integers(2, n + 1))
                mask = sum(1 << int(q) for q in rng.choice(n, size=k, replace=False))
            ops.bitrev(mask)
            U = perm_matrix(n, bitrev_perm(n, mask)) @ U
    o, ud = ops.finish()
    return o, ud, U


def make_dense_small(rng):
    cases = {}
    ci = 0
    for n in range(1, 7):
        for rep in range(6):
            o, ud, U = random_case(rng, n, 12 + 4 * n)
            psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
            psi /= np.linalg.norm(psi)
            out = U @ psi
            cases[f"c{ci}_n"] = np.array(n)
            cases[f"c{ci}_ops"] = o.view(np.uint8)
            cases[f"c{ci}_udata"] = ud
            cases[f"c{ci}_in"] = psi
            cases[f"c{ci}_out"] = out
            ci += 1
    cases["count"] = np.array(ci)
    np.savez_compressed(HERE / "dense_small.npz", **cases)


def make_qft():
    d = {}
    for n in range(2, 13):
        st = normal_state(n, 1000 + n)
        psi = st.to_complex()
        d[f"n{n}_in"] = psi
        d[f"n{n}_out"] = np.fft.ifft(psi) * math.sqrt(1 << n)
    np.savez_compressed(HERE / "qft_fft.npz", **d)


def make_bitrev(rng):
    d = {}
    for n in range(1, 13):
        d[f"full_{n}"] = bitrev_perm(n, 0)
        if n >= 2:
            k = int(rng.integers(2, n + 1))
            mask = sum(1 << int(q) for q in rng.choice(n, size=k, replace=False))
            d[f"mask_{n}"] = np.array(mask)
            d[f"partial_{n}"] = bitrev_perm(n, mask)
    np.savez_compressed(HERE / "bitrev.npz", **d)


def make_grover():
    d = {}
    for n in range(3, 11):
        theta = math.asin(2 ** (-n / 2))
        its = int(math.floor(math.pi / 4 * math.sqrt(2 ** n)))
        d[f"n{n}_p"] = np.array([math.sin((2 * k + 1) * theta) ** 2 for k in range(its + 1)])
        d[f"n{n}_proj"] = np.array([math.cos(2 * k * theta) for k in range(its + 1)])
    np.savez_compressed(HERE / "grover.npz", **d)


def main():
    rng = np.random.default_rng(20261203)
    make_dense_small(rng)
    make_qft()
    make_bitrev(rng)
    make_grover()
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
