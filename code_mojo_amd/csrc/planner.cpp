// planner.cpp -- host-side pass scheduler of the GPU strategy.
//
// The reference runs one CPU kernel per transformation
// (execute_with_handlers, butterfly/core/executors.mojo:268-319); its fastest
// strategy groups consecutive row-local gates and applies a group in one pass
// per row (execute_grid_fused, executors.mojo:1062-1093; grouping rule
// "every target/control < col_bits", executors.mojo:949-955).
//
// Here the rule is generalised to "every non-diagonal target inside the
// tile": a tile is 2^k amplitudes spanned by the low L qubits (for
// coalescing) plus up to k-L further qubits chosen per pass.  Diagonal gates
// (P, CP, Z, CZ, RZ, multi-controlled phases) and all controls may lie
// anywhere, because inside one tile every bit outside the tile is constant.
// Gates are taken out of program order only when they commute with every
// gate left behind, so each pass is exact, not approximate.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bfly_plan.h"

namespace bfly {

namespace {

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

struct Classified {
    int32_t kind;
    bool diag;
    double m[8];
};

// Map a GATE op to the kernel kind the reference's scalar strategy would take
// (transform_scalar / c_transform_scalar / mc_transform_scalar,
// transformations_scalar.mojo:64-175,569-594, with every specialisation on).
Classified classify(const bfly_op& op) {
    Classified c{};
    std::memcpy(c.m, op.mat, sizeof c.m);
    const int nctrl = popc(op.controls);
    const double* g = op.mat;
    const bool mat_diag = g[2] == 0.0 && g[3] == 0.0 && g[4] == 0.0 && g[5] == 0.0;
    auto set_diag = [&](double d0r, double d0i, double d1r, double d1i) {
        c.diag = true;
        if (d0r == 1.0 && d0i == 0.0 && d1r == -1.0 && d1i == 0.0) {
            c.kind = I_NEG1;  // same values as the complex product by -1, exactly
        } else if (d0r == 1.0 && d0i == 0.0) {
            c.kind = I_PHASE1;
        } else {
            c.kind = I_DIAG;
        }
        c.m[0] = d0r; c.m[1] = d0i; c.m[2] = d1r; c.m[3] = d1i;
        for (int i = 4; i < 8; ++i) c.m[i] = 0.0;
    };
    if (nctrl <= 1) {
        if (op.gate == BFLY_GATE_H && nctrl == 0) { c.kind = I_H; return c; }
        if (op.gate == BFLY_GATE_P) { set_diag(1.0, 0.0, std::cos(op.theta), std::sin(op.theta)); return c; }
        if (op.gate == BFLY_GATE_X) { c.kind = I_X; return c; }
        if (op.gate == BFLY_GATE_RY) {
            c.kind = I_RY;
            c.m[0] = std::cos(op.theta / 2);
            c.m[1] = std::sin(op.theta / 2);
            return c;
        }
    }
    if (mat_diag) { set_diag(g[0], g[1], g[6], g[7]); return c; }
    if (g[0] == 0 && g[1] == 0 && g[6] == 0 && g[7] == 0 && g[2] == 1 && g[3] == 0 && g[4] == 1 && g[5] == 0) {
        c.kind = I_X;  // exact exchange, same result as the generic product
        return c;
    }
    c.kind = I_GEN;
    return c;
}

void fill_floats(const double* m, float* mf) {
    for (int i = 0; i < 8; ++i) mf[i] = static_cast<float>(m[i]);
}

// Register form of a classified gate for a tile sweep.  Pair kinds become
// their 2x2 matrix (H with the reference's 0.7071067811865476 entries, X as
// the exchange matrix, RY from cos / sin of theta/2); diagonal kinds keep
// d0 = m[0..1], d1 = m[2..3].
void pair_matrix(const Classified& c, double* out) {
    const double s = 0.7071067811865476;
    switch (c.kind) {
        case I_H: { const double h[8] = {s, 0, s, 0, s, 0, -s, 0}; std::memcpy(out, h, sizeof h); return; }
        case I_X: { const double x[8] = {0, 0, 1, 0, 1, 0, 0, 0}; std::memcpy(out, x, sizeof x); return; }
        case I_RY: {
            const double r[8] = {c.m[0], 0, -c.m[1], 0, c.m[1], 0, c.m[0], 0};
            std::memcpy(out, r, sizeof r);
            return;
        }
        default: std::memcpy(out, c.m, 8 * sizeof(double)); return;
    }
}

uint64_t op_qubits(const bfly_op& op, int n) {
    const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    switch (op.kind) {
        case BFLY_OP_GATE: return (1ull << op.target) | op.controls;
        case BFLY_OP_BITREV: return op.aux ? op.aux : all;
        case BFLY_OP_SWAP: return (1ull << op.target) | (1ull << op.arg_i);
        case BFLY_OP_UNITARY: return (((1ull << op.arg_i) - 1) << op.target) | op.controls;
    }
    return all;
}

GateLaunch make_gate_launch(const bfly_op& op, int n) {
    GateLaunch gl{};
    Classified c = classify(op);
    gl.kind = c.kind;
    gl.target = op.target;
    gl.cmask = op.controls;
    std::memcpy(gl.m, c.m, sizeof gl.m);
    fill_floats(gl.m, gl.mf);
    uint64_t bits = op.controls | (1ull << op.target);
    gl.npos = 0;
    for (int b = 0; b < n; ++b)
        if ((bits >> b) & 1) gl.pos[gl.npos++] = b;
    gl.count = 1ull << (n - gl.npos);
    return gl;
}

struct PendingGate {
    const bfly_op* op;
    Classified c;
};

}  // namespace

int default_tile_bits(int32_t dtype) { return dtype == BFLY_C64 ? 13 : 12; }
int default_low_bits(int32_t dtype) { return dtype == BFLY_C64 ? 5 : 4; }

int validate_ops(int32_t n, const bfly_op* ops, uint64_t nops, uint64_t udata_len, std::string& err) {
    const uint64_t all = (1ull << n) - 1;
    for (uint64_t i = 0; i < nops; ++i) {
        const bfly_op& op = ops[i];
        switch (op.kind) {
            case BFLY_OP_GATE:
                if (op.target < 0 || op.target >= n) { err = "Target out of bounds: " + std::to_string(op.target); return BFLY_EINVAL; }
                if (op.controls & ~all) { err = "Control out of bounds"; return BFLY_EINVAL; }
                if ((op.controls >> op.target) & 1) { err = "Control overlaps target"; return BFLY_EINVAL; }
                if (op.gate < 0 || op.gate > BFLY_GATE_CUSTOM) { err = "Unknown gate kind: " + std::to_string(op.gate); return BFLY_EINVAL; }
                for (int j = 0; j < 8; ++j)
                    if (!std::isfinite(op.mat[j])) { err = "non-finite gate matrix"; return BFLY_EINVAL; }
                break;
            case BFLY_OP_BITREV:
                if (op.aux & ~all) { err = "Target out of bounds"; return BFLY_EINVAL; }
                break;
            case BFLY_OP_SWAP:
                if (op.target < 0 || op.target >= n || op.arg_i < 0 || op.arg_i >= n) { err = "Target out of bounds"; return BFLY_EINVAL; }
                break;
            case BFLY_OP_UNITARY: {
                const int m = op.arg_i;
                if (m <= 0) break;
                if (m > 10) { err = "unitary blocks wider than 10 qubits are not supported"; return BFLY_EUNSUPPORTED; }
                if (op.target < 0 || op.target + m > n) { err = "Target range out of bounds"; return BFLY_EINVAL; }
                const uint64_t range = ((1ull << m) - 1) << op.target;
                if (op.controls & ~all) { err = "Control out of bounds"; return BFLY_EINVAL; }
                if (op.controls & range) { err = "Control overlaps target range"; return BFLY_EINVAL; }
                if (op.aux + (1ull << (2 * m)) > udata_len) { err = "Unitary size mismatch"; return BFLY_EINVAL; }
                break;
            }
            default:
                err = "Unknown transformation type";
                return BFLY_EUNSUPPORTED;
        }
    }
    return BFLY_OK;
}

namespace {

// Pack the gates of one pass into register sweeps and append the records.
// A sweep owns up to kSlots local bits; its gates' control masks are split
// into slot bits / other local bits / outer bits once the slot set is final.
void emit_tile_pass(HostPlan& plan, int n, int k, uint64_t tile_mask, const std::vector<PendingGate>& gates) {
    Step st{};
    st.kind = S_TILE;
    TilePass& tp = st.tile;
    tp.n = n;
    tp.k = k;
    int local_of[64];
    int li = 0, oi = 0;
    for (int b = 0; b < n; ++b) {
        if ((tile_mask >> b) & 1) { local_of[b] = li; tp.qpos[li++] = b; }
        else { local_of[b] = -1; tp.opos[oi++] = b; }
    }
    tp.nouter = oi;
    tp.first_sweep = static_cast<int32_t>(plan.sweeps.size());

    int slots[kSlots];
    int nslots = 0;
    std::vector<const PendingGate*> cur;
    auto slot_of = [&](int lb) {
        for (int s = 0; s < nslots; ++s)
            if (slots[s] == lb) return s;
        return -1;
    };
    auto close = [&]() {
        for (int b = 0; nslots < kSlots && b < k; ++b)
            if (slot_of(b) < 0) slots[nslots++] = b;
        TileSweep sw{};
        for (int s = 0; s < kSlots; ++s) sw.s[s] = slots[s];
        sw.first = static_cast<int32_t>(plan.instrs.size());
        for (const PendingGate* pg : cur) {
            const bfly_op& op = *pg->op;
            TileInstr in{};
            const int ik = pg->c.kind;
            in.kind = ik == I_NEG1 ? T_NEG : (pg->c.diag ? T_DIAG : T_PAIR);
            pair_matrix(pg->c, in.m);
            fill_floats(in.m, in.mf);
            in.dlocal = -1;
            for (int b = 0; b < n; ++b) {
                if (!((op.controls >> b) & 1)) continue;
                const int lb = local_of[b];
                if (lb < 0) in.omask |= 1ull << b;
                else if (slot_of(lb) >= 0) in.cslots |= 1u << slot_of(lb);
                else in.lmask |= 1u << lb;
            }
            const int lt = local_of[op.target];
            in.slot = lt >= 0 ? slot_of(lt) : -1;
            if (pg->c.diag) {
                if (lt < 0) in.obit = 1ull << op.target;
                else if (in.slot < 0) in.dlocal = lt;
                for (uint32_t m = 0; m < static_cast<uint32_t>(kGroup); ++m) {
                    if ((m & in.cslots) != in.cslots) continue;
                    if (in.slot < 0 || ((m >> in.slot) & 1u)) in.mm1 |= static_cast<uint16_t>(1u << m);
                    if (in.slot < 0 || !((m >> in.slot) & 1u)) in.mm0 |= static_cast<uint16_t>(1u << m);
                }
                // only I_DIAG touches target-bit-0 members (P / controlled P / Z leave them as is)
                if (ik != I_DIAG) in.mm0 = 0;
            }
            // Sign flips commute and compose by XOR: a Z-family gate whose
            // selection depends on slot bits only merges into a preceding one.
            const bool uniform_neg = in.kind == T_NEG && in.slot >= 0 && in.lmask == 0 && in.omask == 0;
            if (uniform_neg && static_cast<int32_t>(plan.instrs.size()) > sw.first) {
                TileInstr& prev = plan.instrs.back();
                if (prev.kind == T_NEG && prev.slot >= 0 && prev.lmask == 0 && prev.omask == 0) {
                    prev.mm1 = static_cast<uint16_t>(prev.mm1 ^ in.mm1);
                    continue;
                }
            }
            plan.instrs.push_back(in);
        }
        sw.ninstr = static_cast<int32_t>(plan.instrs.size()) - sw.first;
        plan.sweeps.push_back(sw);
        cur.clear();
        nslots = 0;
    };
    for (const PendingGate& pg : gates) {
        if (!pg.c.diag) {
            const int lt = local_of[pg.op->target];
            if (slot_of(lt) < 0) {
                if (nslots == kSlots) close();
                slots[nslots++] = lt;
            }
        }
        cur.push_back(&pg);
    }
    close();
    tp.nsweeps = static_cast<int32_t>(plan.sweeps.size()) - tp.first_sweep;
    plan.steps.push_back(st);
    plan.passes += 1;
}

}  // namespace

int build_plan(HostPlan& plan, int32_t n, int32_t dtype, const bfly_op* ops, uint64_t nops,
               const double* udata, uint64_t udata_len, const bfly_exec_opts& opts, std::string& err) {
    plan = HostPlan{};
    plan.n = n;
    plan.dtype = dtype;
    plan.opts = opts;
    if (udata && udata_len) plan.udata.assign(udata, udata + 2 * udata_len);
    int rc = validate_ops(n, ops, nops, udata_len, err);
    if (rc) return rc;

    const uint64_t all = (1ull << n) - 1;
    int k = opts.tile_bits > 0 ? opts.tile_bits : default_tile_bits(dtype);
    k = std::min(std::max(k, kMinTileBits), kMaxTileBits);
    const bool fused = opts.mode == BFLY_MODE_FUSED && n >= kMinTileBits;
    if (k > n) k = n;
    int L = opts.low_bits > 0 ? opts.low_bits : default_low_bits(dtype);
    // the tile kernel stages 16-byte vectors: 2 (f64) or 4 (f32) consecutive
    // amplitudes must be consecutive local indices
    L = std::max(L, dtype == BFLY_C64 ? 2 : 1);
    // at least one tile slot stays free for a non-diagonal target, so the
    // first remaining gate always fits and every pass makes progress
    L = std::min(L, k - 1);
    const int max_gates = opts.max_gates_per_pass > 0 ? opts.max_gates_per_pass : (1 << 30);
    const uint64_t low_mask = (1ull << L) - 1;

    // Count gates; emit non-gate ops as their own steps.
    auto emit_standalone = [&](const bfly_op& op) {
        Step st{};
        switch (op.kind) {
            case BFLY_OP_BITREV: {
                uint64_t mask = op.aux;
                if (mask == all) mask = 0;
                if (mask != 0 && popc(mask) <= 1) return;  // nothing to do
                st.kind = S_BITREV; st.mask = mask; break;
            }
            case BFLY_OP_SWAP:
                if (op.target == op.arg_i) return;
                st.kind = S_SWAP; st.a = op.target; st.b = op.arg_i; break;
            case BFLY_OP_UNITARY:
                if (op.arg_i <= 0) return;
                st.kind = S_UNITARY; st.a = op.target; st.b = op.arg_i; st.controls = op.controls; st.uoff = op.aux; break;
            default: return;
        }
        plan.steps.push_back(st);
        plan.passes += 1;
    };

    if (!fused) {
        for (uint64_t i = 0; i < nops; ++i) {
            const bfly_op& op = ops[i];
            if (op.kind == BFLY_OP_GATE) {
                Step st{};
                st.kind = S_GATE;
                st.gate = make_gate_launch(op, n);
                plan.steps.push_back(st);
                plan.passes += 1;
                plan.gates += 1;
            } else {
                emit_standalone(op);
            }
        }
        return BFLY_OK;
    }

    // Fused: greedy commutation-aware pass building.
    std::vector<uint32_t> remaining(nops);
    for (uint64_t i = 0; i < nops; ++i) remaining[i] = static_cast<uint32_t>(i);
    std::vector<Classified> cls(nops);
    for (uint64_t i = 0; i < nops; ++i) {
        if (ops[i].kind == BFLY_OP_GATE) { cls[i] = classify(ops[i]); plan.gates += 1; }
    }
    size_t head = 0;  // remaining[head..] are still to schedule (compacted each pass)
    std::vector<uint32_t> rest;
    std::vector<PendingGate> placed;
    constexpr size_t kScanWindow = 8192;  // gates examined past the first deferral
    while (head < remaining.size()) {
        const bfly_op& front = ops[remaining[head]];
        if (front.kind != BFLY_OP_GATE) {
            emit_standalone(front);
            ++head;
            continue;
        }
        uint64_t tile = low_mask;
        uint64_t d_nd = 0, d_all = 0;  // qubits of deferred gates: non-diagonal targets / all qubits
        placed.clear();
        rest.clear();
        size_t scanned_after_defer = 0;
        bool deferred_any = false;
        size_t idx = head;
        for (; idx < remaining.size(); ++idx) {
            const uint32_t oi = remaining[idx];
            const bfly_op& op = ops[oi];
            if (deferred_any && ++scanned_after_defer > kScanWindow) break;
            if ((d_nd & all) == all) break;  // every qubit blocked: nothing more can move
            if (op.kind != BFLY_OP_GATE) {
                const uint64_t q = op_qubits(op, n);
                d_nd |= q; d_all |= q;
                rest.push_back(oi);
                deferred_any = true;
                continue;
            }
            const Classified& c = cls[oi];
            const uint64_t qall = (1ull << op.target) | op.controls;
            const uint64_t qnd = c.diag ? 0 : (1ull << op.target);
            bool ok = !(qnd & d_all) && !(qall & d_nd) && static_cast<int>(placed.size()) < max_gates;
            if (ok && qnd && !(qnd & tile)) {
                if (popc(tile) < k) tile |= qnd;
                else ok = false;
            }
            if (!ok) {
                d_nd |= qnd; d_all |= qall;
                rest.push_back(oi);
                deferred_any = true;
                continue;
            }
            placed.push_back(PendingGate{&op, c});
        }
        if (placed.empty()) {
            err = "internal planner error: a pass placed no gate";
            return BFLY_EINVAL;
        }
        // complete the tile with the lowest unused qubits (any k-subset works)
        for (int b = 0; popc(tile) < k && b < n; ++b) tile |= 1ull << b;
        emit_tile_pass(plan, n, k, tile, placed);
        // compact: deferred ops keep their order, followed by the unscanned tail
        std::vector<uint32_t> next;
        next.reserve(rest.size() + (remaining.size() - idx));
        next.insert(next.end(), rest.begin(), rest.end());
        next.insert(next.end(), remaining.begin() + static_cast<long>(idx), remaining.end());
        remaining.swap(next);
        head = 0;
    }
    return BFLY_OK;
}

}  // namespace bfly
