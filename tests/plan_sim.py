"""numpy interpreter of a dumped execution plan (test infrastructure).

Replays the schedule that the planner compiled (tile passes with their tile
qubits, sweeps, slots, local / outer control masks; stand-alone gates,
bit reversals, swaps, dense blocks) on a complex128 state in host memory.
Comparing its result with the oracle checks the planner's commutation-based
reordering and its mask translation on the CPU, independently of the GPU
arithmetic."""
from __future__ import annotations

import math

import numpy as np

I_GEN, I_H, I_X, I_RY, I_DIAG, I_PHASE1, I_NEG1 = range(7)
T_PAIR, T_NEG, T_DIAG, T_ROT, T_PH, T_SCALE, T_ROTS, T_ROT3 = range(8)
SQ_HALF = 0.7071067811865476


def _parse_m(tok):
    return [float.fromhex(x) for x in tok.split("=", 1)[1].split(",")]


def _kv(parts):
    out = {}
    for p in parts:
        if "=" in p:
            k, v = p.split("=", 1)
            out[k] = v
    return out


def parse(text: str):
    lines = [ln for ln in text.strip().split("\n") if ln]
    head = _kv(lines[0].split()[1:])
    steps = []
    cur_tile = None
    cur_sweep = None
    for ln in lines[1:]:
        parts = ln.split()
        kind = parts[0]
        kv = _kv(parts[1:])
        if kind == "tile":
            cur_tile = {"kind": "tile", "k": int(kv["k"]), "slots": int(kv["slots"]),
                        "qpos": [int(x) for x in kv["qpos"].split(",")], "sweeps": []}
            steps.append(cur_tile)
        elif kind == "sweep":
            cur_sweep = {"s": [int(x) for x in kv["s"].split(",")][:cur_tile["slots"]], "instrs": []}
            cur_tile["sweeps"].append(cur_sweep)
        elif kind == "instr":
            cur_sweep["instrs"].append({"kind": int(kv["kind"]), "slot": int(kv["slot"]), "cslots": int(kv["cslots"]),
                                        "lmask": int(kv["lmask"]), "lval": int(kv["lval"]), "dlocal": int(kv["dlocal"]),
                                        "mm1": int(kv["mm1"]), "mm0": int(kv["mm0"]),
                                        "omask": int(kv["omask"]), "oval": int(kv["oval"]), "obit": int(kv["obit"]),
                                        "m": _parse_m(parts[-1])})
        elif kind == "gate":
            steps.append({"kind": "gate", "gkind": int(kv["kind"]), "target": int(kv["target"]),
                          "cmask": int(kv["cmask"]), "m": _parse_m(parts[-1])})
        elif kind == "bitrev":
            steps.append({"kind": "bitrev", "mask": int(kv["mask"])})
        elif kind == "swap":
            steps.append({"kind": "swap", "a": int(kv["a"]), "b": int(kv["b"])})
        elif kind == "block":
            cur_tile = None
            steps.append({"kind": "block", "k": int(kv["k"]), "renorm": int(kv["renorm"]),
                          "qloc": [int(x) for x in kv["qloc"].split(",")],
                          "qpos": [int(x) for x in kv["qpos"].split(",")],
                          "mat": np.array([float.fromhex(x) for x in kv["mat"].split(",")]).reshape(64, 64),
                          "pre": [], "post": []})
        elif kind == "phase":
            steps[-1][kv["when"]].append((int(kv["omask"]), int(kv["slot"]), float.fromhex(kv["theta"])))
        elif kind == "norm":
            steps.append({"kind": "norm"})
        elif kind == "flip":
            steps.append({"kind": "flip", "mask": int(kv["mask"]), "value": int(kv["value"])})
        elif kind == "diffuse":
            steps.append({"kind": "diffuse", "mask": int(kv["mask"]), "variant": int(kv["variant"]),
                          "chain": int(kv.get("chain", 0)), "fmask": int(kv.get("fmask", 0)),
                          "fvalue": int(kv.get("fvalue", 0))})
        elif kind == "unitary":
            steps.append({"kind": "unitary", "target": int(kv["target"]), "m": int(kv["m"]),
                          "controls": int(kv["controls"]), "uoff": int(kv["uoff"])})
        elif kind == "measure":
            pairs = [tuple(int(x) for x in tv.split(":")) for tv in kv["targets"].split(",")]
            steps.append({"kind": "measure", "seeded": int(kv["seeded"]), "seed": int(kv["seed"]),
                          "targets": [p[0] for p in pairs], "values": [p[1] for p in pairs]})
    return int(head["n"]), steps


def apply_measure(psi: np.ndarray, n: int, targets, values, seed: int) -> np.ndarray:
    """apply_measure (butterfly/core/state.mojo:447-553) in FP64 numpy: outcome
    probabilities of the free targets among the amplitudes matching the forced
    values (-1 = free), one MINSTD step from `seed`, the first outcome whose
    running sum reaches r, collapse and division by sqrt(p)."""
    idx = np.arange(1 << n, dtype=np.int64)
    match = np.ones(1 << n, dtype=bool)
    free = []
    for t, v in zip(targets, values):
        if v >= 0:
            match &= ((idx >> t) & 1) == v
        else:
            free.append(t)
    outcome = np.zeros(1 << n, dtype=np.int64)
    for k, t in enumerate(free):
        outcome |= ((idx >> t) & 1) << k
    p = np.abs(psi) ** 2
    probs = np.zeros(1 << len(free))
    np.add.at(probs, outcome[match], p[match])
    total = probs.sum()
    if total <= 0:
        raise ValueError("Measurement probability is zero")
    rng = (seed % 2147483647) * 48271 % 2147483647
    r = rng / 2147483647 * total
    cum = np.cumsum(probs)
    hit = np.nonzero(r <= cum)[0]
    chosen = int(hit[0]) if hit.size else 0
    if probs[chosen] <= 0:
        raise ValueError("Measurement outcome probability is zero")
    keep = match & (outcome == chosen)
    return np.where(keep, psi / np.sqrt(probs[chosen]), 0)


def _apply_pair(kind, m, u, v):
    if kind == I_H:
        return (u + v) * SQ_HALF, (u - v) * SQ_HALF
    if kind == I_X:
        return v, u
    if kind == I_RY:
        c, s = m[0], m[1]
        return u * c - v * s, u * s + v * c
    if kind == I_DIAG:
        return u * complex(m[0], m[1]), v * complex(m[2], m[3])
    if kind == I_PHASE1:
        return u, v * complex(m[2], m[3])
    if kind == I_NEG1:
        return u, -v
    g = [complex(m[0], m[1]), complex(m[2], m[3]), complex(m[4], m[5]), complex(m[6], m[7])]
    return g[0] * u + g[1] * v, g[2] * u + g[3] * v


def run_plan(text: str, psi: np.ndarray, udata: np.ndarray) -> np.ndarray:
    n, steps = parse(text)
    psi = psi.astype(np.complex128).copy()
    idx = np.arange(1 << n, dtype=np.int64)
    for st in steps:
        if st["kind"] == "tile":
            k, qpos = st["k"], st["qpos"]
            outer = [b for b in range(n) if b not in qpos]
            ntiles = 1 << (n - k)
            t = np.arange(ntiles, dtype=np.int64)
            base = np.zeros(ntiles, dtype=np.int64)
            for i, b in enumerate(outer):
                base |= ((t >> i) & 1) << b
            j = np.arange(1 << k, dtype=np.int64)
            l2g = np.zeros(1 << k, dtype=np.int64)
            for i, b in enumerate(qpos):
                l2g |= ((j >> i) & 1) << b
            G = base[:, None] + l2g[None, :]
            A = psi[G]
            for sw in st["sweeps"]:
                S = st["slots"]
                assert S in (3, 4) and len(set(sw["s"])) == S and all(0 <= x < k for x in sw["s"])
                for ins in sw["instrs"]:
                    tile_ok = (base & ins["omask"]) == ins["oval"]
                    lane_ok = (j & ins["lmask"]) == ins["lval"]
                    # member index of every local index within its 16-group
                    member = np.zeros(1 << k, dtype=np.int64)
                    for si in range(S):
                        member |= ((j >> sw["s"][si]) & 1) << si
                    m = ins["m"]
                    if ins["kind"] == T_ROT3:
                        assert S == 3 and ins["cslots"] == 7
                        assert all(math.copysign(1.0, m[3 * sl]) > 0 for sl in range(3))
                    if ins["kind"] in (T_ROTS, T_ROT3):
                        for sl in range(S):
                            if not (ins["cslots"] >> sl) & 1:
                                continue
                            tf, zr, zi = m[3 * sl], m[3 * sl + 1], m[3 * sl + 2]
                            lb = sw["s"][sl]
                            j0 = j[((j >> lb) & 1) == 0]
                            j1 = j0 | (1 << lb)
                            u, v = A[:, j0], A[:, j1] * complex(zr, zi)
                            t = abs(tf)
                            if math.copysign(1.0, tf) < 0:
                                u, v = t * u - v, u + t * v
                            else:
                                u, v = u - t * v, v + t * u
                            A[:, j0], A[:, j1] = u, v
                        continue
                    if ins["kind"] in (T_ROT, T_PH, T_SCALE):
                        # factored one-qubit unitaries: never controlled
                        assert ins["lmask"] == 0 and ins["omask"] == 0 and ins["cslots"] == 0
                        if ins["kind"] == T_SCALE:
                            A = A * complex(m[0], m[1])
                            continue
                        lb = sw["s"][ins["slot"]]
                        j0 = j[((j >> lb) & 1) == 0]
                        j1 = j0 | (1 << lb)
                        u, v = A[:, j0], A[:, j1]
                        if ins["kind"] == T_PH:
                            A[:, j1] = v * complex(m[0], m[1])
                        else:
                            v = v * complex(m[2], m[3])   # pre-phase
                            if m[1]:
                                u, v = m[0] * u - v, u + m[0] * v
                            else:
                                u, v = u - m[0] * v, v + m[0] * u
                            A[:, j0], A[:, j1] = u, v
                    elif ins["kind"] == T_PAIR:
                        cs = ins["cslots"]
                        lb = sw["s"][ins["slot"]]
                        j0 = j[((j >> lb) & 1) == 0]
                        j1 = j0 | (1 << lb)
                        on = ((member[j0] & cs) == cs) & lane_ok[j0]
                        cond = tile_ok[:, None] & on[None, :]
                        u, v = A[:, j0], A[:, j1]
                        nu, nv = _apply_pair(I_GEN, m, u, v)
                        A[:, j0] = np.where(cond, nu, u)
                        A[:, j1] = np.where(cond, nv, v)
                    else:
                        sel1 = (ins["mm1"] >> member) & 1
                        sel0 = (ins["mm0"] >> member) & 1
                        if ins["slot"] < 0:
                            if ins["dlocal"] >= 0:
                                bit = ((j >> ins["dlocal"]) & 1)[None, :] * np.ones((ntiles, 1), np.int64)
                            else:
                                bit = ((base & ins["obit"]) != 0).astype(np.int64)[:, None] * np.ones((1, 1 << k), np.int64)
                            if ins["kind"] == T_NEG:
                                # sign flips: mm1 where the target bit is 1, mm0 where it is 0
                                s1 = sel1[None, :] * bit + sel0[None, :] * (1 - bit)
                                s0 = 0 * s1
                            else:
                                s1 = sel1[None, :] * bit
                                s0 = sel0[None, :] * (1 - bit)
                        else:
                            s1 = sel1[None, :] * np.ones((ntiles, 1), np.int64)
                            s0 = sel0[None, :] * np.ones((ntiles, 1), np.int64)
                        active = tile_ok[:, None] & lane_ok[None, :]
                        if ins["kind"] == T_NEG and ins["slot"] >= 0 and S == 3:
                            # 3-slot kernels multiply by the record's member factors
                            f = np.asarray(m[:8])[member][None, :] * np.ones((ntiles, 1))
                        elif ins["kind"] == T_NEG:
                            f = np.where(s1 == 1, -1.0, 1.0)
                        else:
                            f = np.where(s1 == 1, complex(m[2], m[3]), np.where(s0 == 1, complex(m[0], m[1]), 1.0))
                        A = np.where(active, A * f, A)
            psi[G] = A
        elif st["kind"] == "gate":
            t = st["target"]
            cm = st["cmask"]
            i0 = idx[(((idx >> t) & 1) == 0) & ((idx & cm) == cm)]
            i1 = i0 | (1 << t)
            nu, nv = _apply_pair(st["gkind"], st["m"], psi[i0], psi[i1])
            psi[i0], psi[i1] = nu, nv
        elif st["kind"] == "bitrev":
            mask = st["mask"] or ((1 << n) - 1)
            pos = [b for b in range(n) if (mask >> b) & 1]
            perm = idx & ~mask
            for i, b in enumerate(pos):
                perm |= ((idx >> b) & 1) << pos[len(pos) - 1 - i]
            psi = psi[perm]
        elif st["kind"] == "block":
            qbits = [st["qpos"][l] for l in st["qloc"]]
            if st["renorm"]:
                psi = psi / np.sqrt(np.vdot(psi, psi).real)

            qmask = sum(1 << q for q in qbits)

            def phases(lst):
                # a record's omask is tested on the column index (slot bits
                # cleared), as the block kernel does
                ang = np.zeros(1 << n)
                for om, slot, th in lst:
                    assert om & qmask == 0, "phase record names a block qubit"
                    hit = (idx & ~qmask & om) == om
                    if slot >= 0:
                        hit &= ((idx >> qbits[slot]) & 1) == 1
                    ang += np.where(hit, th, 0.0)
                return np.exp(1j * ang)

            psi = psi * phases(st["pre"])
            m = st["mat"]
            U = m[:32, :32] + 1j * m[32:, :32]
            bases = idx[(idx & qmask) == 0]
            e = np.arange(32)
            offs = np.zeros(32, dtype=np.int64)
            for jslot, q in enumerate(qbits):
                offs |= ((e >> jslot) & 1) << q
            cols = bases[:, None] + offs[None, :]
            psi[cols] = psi[cols] @ U.T
            psi = psi * phases(st["post"])
        elif st["kind"] == "norm":
            pass
        elif st["kind"] == "flip":
            psi = np.where((idx & st["mask"]) == st["value"], -psi, psi)
        elif st["kind"] == "diffuse":
            mask = st["mask"]
            others = ((1 << n) - 1) & ~mask
            key = idx & others
            sums = np.zeros(1 << n, dtype=np.complex128)
            np.add.at(sums, key, psi)
            mean = sums[key] / (1 << bin(mask).count("1"))
            if st["variant"] == 0:
                psi = psi - 2 * mean
            else:
                psi = (2 * mean.real - psi.real) + 1j * (-2 * mean.imag - psi.imag)
            if st["chain"] & 8:   # the phase flip absorbed into this diffuser's pass
                psi = np.where((idx & st["fmask"]) == st["fvalue"], -psi, psi)
        elif st["kind"] == "swap":
            a, b = st["a"], st["b"]
            x = ((idx >> a) & 1) ^ ((idx >> b) & 1)
            psi = psi[idx ^ ((x << a) | (x << b))]
        elif st["kind"] == "unitary":
            m, t, cm = st["m"], st["target"], st["controls"]
            dim = 1 << m
            u = udata[2 * st["uoff"]: 2 * (st["uoff"] + dim * dim)]
            U = (u[0::2] + 1j * u[1::2]).reshape(dim, dim)
            block = ((1 << m) - 1) << t
            bases = idx[((idx & block) == 0) & ((idx & cm) == cm)]
            cols = bases[:, None] + (np.arange(dim)[None, :] << t)
            psi[cols] = psi[cols] @ U.T
        elif st["kind"] == "measure":
            if not st["seeded"]:
                raise ValueError("replay needs a measurement seed")
            psi = apply_measure(psi, n, st["targets"], st["values"], st["seed"])
    return psi
