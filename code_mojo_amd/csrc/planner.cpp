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
#include <cstdlib>
#include <cmath>
#include <complex>
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

void fill_floats(const double* m, float* mf, int count = 8) {
    for (int i = 0; i < count; ++i) mf[i] = static_cast<float>(m[i]);
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
        case BFLY_OP_PHASE_FLIP:
        case BFLY_OP_DIFFUSE: return op.aux;
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
// 3-slot sweeps (8 amplitudes per thread) from 2^10-amplitude tiles for both
// precisions: twice the threads of 4-slot sweeps and straight-line three-slot
// rotation records.  For complex64 2^13 tiles (1024 threads, two blocks per
// CU) the kernel spills 28 bytes per thread (the tile's global addressing,
// stored once per tile and reloaded for its stage-out, none in the gate loop)
// and still runs 8 % faster than with 4 slots (30q layered 435 vs 471 ms);
// one block per CU without spills is slower (560 ms).
int default_sweep_slots(int32_t dtype, int k) { (void)dtype; return k >= kMinSlots3Bits ? 3 : 4; }
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
            case BFLY_OP_PHASE_FLIP:
                if (op.aux == 0 || (op.aux & ~all)) { err = "Target out of bounds"; return BFLY_EINVAL; }
                if (op.controls & ~op.aux) { err = "phase-flip pattern outside its mask"; return BFLY_EINVAL; }
                break;
            case BFLY_OP_MEASURE: {
                if (op.arg_i < 0 || op.aux + static_cast<uint64_t>(op.arg_i) > udata_len) {
                    err = "Measurement values size mismatch";
                    return BFLY_EINVAL;
                }
                if (op.gate != 0 && op.gate != 1) { err = "bad measurement seed flag"; return BFLY_EINVAL; }
                if (op.gate == 1 && static_cast<int64_t>(op.controls) < 0) { err = "negative measurement seed"; return BFLY_EINVAL; }
                break;
            }
            case BFLY_OP_DIFFUSE:
                if (op.aux == 0 || (op.aux & ~all)) { err = "Target out of bounds"; return BFLY_EINVAL; }
                if (op.arg_i != 0 && op.arg_i != 1) { err = "unknown diffuser variant"; return BFLY_EINVAL; }
                break;
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
struct Cx {
    double r, i;
};
Cx cmul(Cx a, Cx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
Cx cdiv_unit(Cx a, Cx b) { return {a.r * b.r + a.i * b.i, a.i * b.r - a.r * b.i}; }  // a / b, |b| = 1
bool is_one(Cx z) { return z.r == 1.0 && z.i == 0.0; }
bool is_minus_one(Cx z) { return z.r == -1.0 && z.i == 0.0; }

// Factors a 2x2 unitary as M = sigma * PH(ephi) * R * PH(elam), where
// PH(z) = diag(1, z) and R is the real rotation [[c, -s], [s, c]] written as
// c * [[1, -t], [t, 1]] (form 0, t = s/c <= 2^20) or s * [[t, -1], [1, t]]
// (form 1, t = c/s < 2^-20: rotations within about 1e-6 of a quarter turn,
// such as X and Y); sigma collects e^{i alpha} and c or s.  A large t costs
// no accuracy: each output is one FMA rounded relative to its own size, and
// the pass's scale product is applied before it falls below 2^-16.  Keeping
// almost every rotation in form 0 lets the kernels run them without
// branches.  Returns false when M is not unitary to 1e-12 (then the generic
// pair product runs).
bool factor_unitary(const double* m, Cx& sigma, Cx& ephi, Cx& elam, int& form, double& t) {
    const Cx m00{m[0], m[1]}, m01{m[2], m[3]}, m10{m[4], m[5]}, m11{m[6], m[7]};
    const double c = std::hypot(m00.r, m00.i), s = std::hypot(m10.r, m10.i);
    if (!(c + s > 0.0) || !std::isfinite(c + s)) return false;
    const Cx ea = c > 0.0 ? Cx{m00.r / c, m00.i / c} : Cx{m10.r / s, m10.i / s};
    ephi = {1.0, 0.0};
    if (c > 0.0 && s > 0.0) {
        ephi = cdiv_unit({m10.r / s, m10.i / s}, ea);
        elam = cdiv_unit({-m01.r / s, -m01.i / s}, ea);
    } else if (s == 0.0) {
        elam = cdiv_unit({m11.r / c, m11.i / c}, ea);
    } else {
        elam = cdiv_unit({-m01.r / s, -m01.i / s}, ea);
    }
    double scale;
    if (c >= s * 0x1p-20) { form = 0; t = s / c; scale = c; }
    else { form = 1; t = c / s; scale = s; }
    sigma = {ea.r * scale, ea.i * scale};
    // rebuild ea * [[c, -elam s], [ephi s, ephi elam c]] and compare
    const Cx el_s = cmul(elam, {s, 0.0}), ep_s = cmul(ephi, {s, 0.0});
    const Cx rec[4] = {cmul(ea, {c, 0.0}), cmul(ea, {-el_s.r, -el_s.i}), cmul(ea, ep_s),
                       cmul(ea, cmul(cmul(ephi, elam), {c, 0.0}))};
    const Cx ref[4] = {m00, m01, m10, m11};
    for (int e = 0; e < 4; ++e)
        if (std::fabs(rec[e].r - ref[e].r) > 1e-12 || std::fabs(rec[e].i - ref[e].i) > 1e-12) return false;
    return std::fabs(c * c + s * s - 1.0) <= 1e-12;
}

// One entry of a tile pass's instruction list before sweep assignment.
struct Micro {
    enum Kind { GATE, ROT, PH, SCALE } kind;
    int q;                    // target qubit (global); -1 for SCALE
    const PendingGate* pg;    // GATE
    Cx z;                     // PH factor on target bit 1; SCALE factor
    int form;                 // ROT
    double t;                 // ROT
};

bool blocks_phase(const Micro& u, int q) {
    return u.q == q && (u.kind == Micro::ROT || (u.kind == Micro::GATE && !u.pg->c.diag));
}

// Phases (factor on target bit 1) carried from one tile pass to a later one.
using PhaseCarry = std::vector<std::pair<int, Cx>>;

// Builds the instruction list of a tile pass: uncontrolled one-qubit unitaries
// are factored (T_ROT between phases), phases on one qubit are moved forward
// through everything that commutes with them (diagonal gates, gates on other
// targets, gates controlled by that qubit) and multiplied together, and the
// pass's scale factors are applied by T_SCALE whenever their product falls
// below 2^-16 and at the end.  Phases carried in from earlier passes go
// first; a phase with no non-diagonal action on its qubit after it in this
// pass leaves through `carry_out` when `may_carry(q)` (a later pass applies
// it, usually as the pre-phase of the qubit's next rotation, which costs
// nothing) instead of becoming an instruction of its own.
template <typename MayCarry>
std::vector<Micro> micro_ops(const std::vector<PendingGate>& gates, bool factor, const PhaseCarry& carry_in,
                             PhaseCarry* carry_out, MayCarry&& may_carry) {
    std::vector<Micro> u;
    u.reserve(gates.size() * 3 + carry_in.size() + 4);
    for (const auto& c : carry_in) u.push_back({Micro::PH, c.first, nullptr, c.second, 0, 0.0});
    for (const PendingGate& pg : gates) {
        const bfly_op& op = *pg.op;
        const bool gate = op.kind == BFLY_OP_GATE && factor;
        if (gate && op.controls == 0 && pg.c.diag && (pg.c.kind == I_PHASE1 || pg.c.kind == I_NEG1)) {
            const Cx z = pg.c.kind == I_NEG1 ? Cx{-1.0, 0.0} : Cx{pg.c.m[2], pg.c.m[3]};
            u.push_back({Micro::PH, op.target, nullptr, z, 0, 0.0});
            continue;
        }
        if (gate && op.controls == 0 && !pg.c.diag) {
            double m[8];
            pair_matrix(pg.c, m);
            Cx sigma, ephi, elam;
            int form;
            double t;
            if (factor_unitary(m, sigma, ephi, elam, form, t)) {
                if (!is_one(elam)) u.push_back({Micro::PH, op.target, nullptr, elam, 0, 0.0});
                u.push_back({Micro::ROT, op.target, nullptr, sigma, form, t});
                if (!is_one(ephi)) u.push_back({Micro::PH, op.target, nullptr, ephi, 0, 0.0});
                continue;
            }
        }
        u.push_back({Micro::GATE, op.kind == BFLY_OP_GATE ? op.target : -1, &pg, {1.0, 0.0}, 0, 0.0});
    }
    const size_t N = u.size();
    std::vector<char> live(N, 1);
    // merge each phase into the next phase on its qubit, unless a
    // non-diagonal action on that qubit comes first
    for (size_t i = 0; i < N; ++i) {
        if (u[i].kind != Micro::PH) continue;
        for (size_t j = i + 1; j < N; ++j) {
            if (!live[j]) continue;
            if (u[j].kind == Micro::PH && u[j].q == u[i].q) {
                u[j].z = cmul(u[j].z, u[i].z);
                live[i] = 0;
                break;
            }
            if (blocks_phase(u[j], u[i].q)) break;
        }
    }
    // anchor every remaining phase next to a non-diagonal action on its
    // qubit (just before the next one, else just after the previous one), so
    // that it runs in a sweep where its qubit is a slot
    std::vector<std::pair<long, size_t>> order;
    order.reserve(N);
    for (size_t i = 0; i < N; ++i) {
        if (!live[i]) continue;
        long key = 2 * static_cast<long>(i);
        if (u[i].kind == Micro::PH) {
            if (is_one(u[i].z)) continue;
            long anchor = -1;
            for (size_t j = i + 1; j < N && anchor < 0; ++j)
                if (live[j] && blocks_phase(u[j], u[i].q)) anchor = 2 * static_cast<long>(j) - 1;
            if (anchor < 0 && carry_out && may_carry(u[i].q)) {
                carry_out->push_back({u[i].q, u[i].z});
                continue;
            }
            for (size_t j = i; j-- > 0 && anchor < 0;)
                if (live[j] && blocks_phase(u[j], u[i].q)) anchor = 2 * static_cast<long>(j) + 1;
            if (anchor >= 0) key = anchor;
        }
        order.push_back({key, i});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<long, size_t>& a, const std::pair<long, size_t>& b) { return a.first < b.first; });
    std::vector<Micro> out;
    out.reserve(order.size() + 4);
    Cx acc{1.0, 0.0};
    for (const auto& e : order) {
        Micro m = u[e.second];
        out.push_back(m);
        if (m.kind == Micro::ROT) {
            acc = cmul(acc, m.z);
            if (std::hypot(acc.r, acc.i) < 0x1p-16) {
                out.push_back({Micro::SCALE, -1, nullptr, acc, 0, 0.0});
                acc = {1.0, 0.0};
            }
        }
    }
    if (!is_one(acc)) out.push_back({Micro::SCALE, -1, nullptr, acc, 0, 0.0});
    return out;
}

// -- small dense unitaries as gates ---------------------------------------
// Fused mode runs a UNITARY op as a pass of its own (unitary_kernel).  The
// reference's fusion rewrites emit two kinds that are plain gates in
// disguise (transformations_fusion.mojo:334-384,539-660): m = 1 blocks (a run
// of gates on one target, with or without one control) and m = 2 blocks of
// uncontrolled gates on adjacent qubits, which are Kronecker products
// G_hi (x) G_lo.  Both become 2x2 gate records that join the tile passes: an
// m = 1 block directly, an m = 2 block when it factors to 1e-14 (a controlled
// G_hi (x) G_lo is controlled-G_hi times controlled-G_lo, so the controls carry
// over).  Other blocks stay stand-alone steps.
bfly_op gate_from_2x2(const std::complex<double>* g, int target, uint64_t controls) {
    bfly_op o{};
    o.kind = BFLY_OP_GATE;
    o.gate = BFLY_GATE_CUSTOM;
    o.target = target;
    o.controls = controls;
    for (int i = 0; i < 4; ++i) {
        o.mat[2 * i] = g[i].real();
        o.mat[2 * i + 1] = g[i].imag();
    }
    return o;
}

// U (4 x 4, row-major, index b0 + 2 b1) = A (x) B with A on b1, B on b0?
bool factor_kron(const std::complex<double>* U, std::complex<double>* A, std::complex<double>* B) {
    using C = std::complex<double>;
    auto blk = [&](int i1, int j1, int r0, int c0) { return U[(r0 + 2 * i1) * 4 + (c0 + 2 * j1)]; };
    int bi = 0, bj = 0;
    double best = -1.0;
    for (int i1 = 0; i1 < 2; ++i1)
        for (int j1 = 0; j1 < 2; ++j1) {
            double f = 0;
            for (int r = 0; r < 2; ++r)
                for (int c = 0; c < 2; ++c) f += std::norm(blk(i1, j1, r, c));
            if (f > best) { best = f; bi = i1; bj = j1; }
        }
    if (!(best > 0)) return false;
    const double scale = std::sqrt(2.0 / best);   // a unitary 2x2 has Frobenius norm sqrt(2)
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) B[r * 2 + c] = blk(bi, bj, r, c) * scale;
    for (int i1 = 0; i1 < 2; ++i1)
        for (int j1 = 0; j1 < 2; ++j1) {
            C t = 0;   // tr(B^H U_blk) / 2
            for (int r = 0; r < 2; ++r)
                for (int c = 0; c < 2; ++c) t += std::conj(B[r * 2 + c]) * blk(i1, j1, r, c);
            A[i1 * 2 + j1] = t * 0.5;
        }
    for (int i1 = 0; i1 < 2; ++i1)
        for (int j1 = 0; j1 < 2; ++j1)
            for (int r = 0; r < 2; ++r)
                for (int c = 0; c < 2; ++c)
                    if (std::abs(A[i1 * 2 + j1] * B[r * 2 + c] - blk(i1, j1, r, c)) > 1e-14) return false;
    return true;
}

std::vector<bfly_op> unitaries_as_gates(const bfly_op* ops, uint64_t nops, const double* udata, bool& changed) {
    std::vector<bfly_op> out;
    out.reserve(nops + 8);
    changed = false;
    for (uint64_t i = 0; i < nops; ++i) {
        const bfly_op& op = ops[i];
        if (op.kind == BFLY_OP_UNITARY && (op.arg_i == 1 || op.arg_i == 2)) {
            const auto* u = reinterpret_cast<const std::complex<double>*>(udata + 2 * op.aux);
            if (op.arg_i == 1) {
                out.push_back(gate_from_2x2(u, op.target, op.controls));
                changed = true;
                continue;
            }
            std::complex<double> A[4], B[4];
            if (factor_kron(u, A, B)) {
                out.push_back(gate_from_2x2(B, op.target, op.controls));
                out.push_back(gate_from_2x2(A, op.target + 1, op.controls));
                changed = true;
                continue;
            }
        }
        out.push_back(op);
    }
    return out;
}

// -- runs of gates on one target -------------------------------------------
// Adjacent GATE ops with the same target and the same controls act as one
// 2x2 gate: their matrices multiply (later gates on the left) in FP64, each
// taken as the kernels apply it (H with the reference's 0.7071067811865476
// entries, P and RY from cos / sin of the angle, X as the exchange).  Runs are
// cut at 64 gates, which keeps the product's rounding (about one unit in the
// last place per factor) far inside the 1e-12 amplitude bar.  Amplitude
// estimation's controlled powers of its Grover operator are such runs: the
// 9-qubit circuit of 1569 gates, all on qubit 0, becomes 44 gates.
void gate_matrix(const bfly_op& op, std::complex<double>* g) {
    using C = std::complex<double>;
    const Classified c = classify(op);
    switch (c.kind) {
        case I_H: {
            const double h = 0.7071067811865476;
            g[0] = C(h, 0); g[1] = C(h, 0); g[2] = C(h, 0); g[3] = C(-h, 0);
            return;
        }
        case I_X:
            g[0] = C(0, 0); g[1] = C(1, 0); g[2] = C(1, 0); g[3] = C(0, 0);
            return;
        case I_RY:
            g[0] = C(c.m[0], 0); g[1] = C(-c.m[1], 0); g[2] = C(c.m[1], 0); g[3] = C(c.m[0], 0);
            return;
        case I_NEG1:
        case I_PHASE1:
        case I_DIAG:
            g[0] = C(c.m[0], c.m[1]); g[1] = C(0, 0); g[2] = C(0, 0); g[3] = C(c.m[2], c.m[3]);
            return;
        default:
            for (int i = 0; i < 4; ++i) g[i] = C(c.m[2 * i], c.m[2 * i + 1]);
    }
}

std::vector<bfly_op> merge_gate_runs(const bfly_op* ops, uint64_t nops, bool& changed) {
    using C = std::complex<double>;
    constexpr int kMaxRun = 64;
    std::vector<bfly_op> out;
    out.reserve(nops);
    changed = false;
    uint64_t i = 0;
    while (i < nops) {
        const bfly_op& op = ops[i];
        uint64_t j = i + 1;
        if (op.kind == BFLY_OP_GATE) {
            while (j < nops && j - i < static_cast<uint64_t>(kMaxRun) && ops[j].kind == BFLY_OP_GATE &&
                   ops[j].target == op.target && ops[j].controls == op.controls)
                ++j;
        }
        if (j - i == 1) {
            out.push_back(op);
        } else {
            C m[4], g[4], r[4];
            gate_matrix(ops[i], m);
            for (uint64_t q = i + 1; q < j; ++q) {
                gate_matrix(ops[q], g);
                r[0] = g[0] * m[0] + g[1] * m[2];
                r[1] = g[0] * m[1] + g[1] * m[3];
                r[2] = g[2] * m[0] + g[3] * m[2];
                r[3] = g[2] * m[1] + g[3] * m[3];
                for (int e = 0; e < 4; ++e) m[e] = r[e];
            }
            out.push_back(gate_from_2x2(m, op.target, op.controls));
            changed = true;
        }
        i = j;
    }
    return out;
}

// -- Grover peephole -------------------------------------------------------
// Recognises, in a gate list, the sequences that the reference's classical
// replacement rules recognise (grover_diffuser_replacement_rule /
// grover_inversion_0_replacement_rule, core/quantum_circuit.mojo:298-495) in
// a more general form, and replaces them with their exact action:
//   X on S, multi-controlled phase -1 on T (S within T), X on S
//       -> PHASE_FLIP(mask T, pattern T \ S)   (phase_oracle_match, inversion_0)
//   H on T, PHASE_FLIP(T, 0), H on T
//       -> DIFFUSE(T), a <- a - 2 mean_T(a)      (diffuser_circuit)
bool is_plain(const bfly_op& op, int gate) {
    return op.kind == BFLY_OP_GATE && op.controls == 0 && op.gate == gate;
}

// Run of gates `gate` without controls on distinct targets starting at i.
int plain_run(const bfly_op* ops, uint64_t nops, uint64_t i, int gate, uint64_t& set) {
    set = 0;
    int len = 0;
    while (i + len < nops && is_plain(ops[i + len], gate) && !((set >> ops[i + len].target) & 1)) {
        set |= 1ull << ops[i + len].target;
        ++len;
    }
    return len;
}

bool is_sign_gate(const bfly_op& op) {
    if (op.kind != BFLY_OP_GATE) return false;
    const Classified c = classify(op);
    if (!c.diag || c.m[0] != 1.0 || c.m[1] != 0.0) return false;
    return std::fabs(c.m[2] + 1.0) <= 1e-15 && std::fabs(c.m[3]) <= 1e-15;
}

// X-conjugated sign gate at i: returns its length (0 if none), mask and pattern.
int match_flip(const bfly_op* ops, uint64_t nops, uint64_t i, int n, uint64_t& mask, uint64_t& value) {
    uint64_t S = 0;
    const int s = plain_run(ops, nops, i, BFLY_GATE_X, S);
    if (i + s >= nops || !is_sign_gate(ops[i + s])) return 0;
    const bfly_op& g = ops[i + s];
    const uint64_t T = g.controls | (1ull << g.target);
    if (S & ~T) return 0;
    if (s == 0 && popc(T) < n) return 0;  // a bare controlled Z stays a gate
    uint64_t S2 = 0;
    const int s2 = plain_run(ops, nops, i + s + 1, BFLY_GATE_X, S2);
    if (s2 < s || S2 != S) {
        // the closing run may be longer than S: only its first s gates belong here
        uint64_t acc = 0;
        for (int j = 0; j < s; ++j) {
            const uint64_t q = i + s + 1 + j;
            if (q >= nops || !is_plain(ops[q], BFLY_GATE_X)) return 0;
            acc |= 1ull << ops[q].target;
        }
        if (acc != S) return 0;
    }
    mask = T;
    value = T & ~S;
    return 2 * s + 1;
}

std::vector<bfly_op> rewrite_grover(const bfly_op* ops, uint64_t nops, int n, bool diffusers) {
    std::vector<bfly_op> out;
    out.reserve(nops);
    uint64_t i = 0;
    while (i < nops) {
        uint64_t mask = 0, value = 0;
        if (diffusers) {
            uint64_t T = 0;
            const int t = plain_run(ops, nops, i, BFLY_GATE_H, T);
            int f = 0;
            if (t > 0 && (popc(T) <= kDiffuseMaxBits || popc(T) == n))
                f = match_flip(ops, nops, i + t, n, mask, value);
            if (f > 0 && mask == T && value == 0 && f == 2 * t + 1) {
                uint64_t T2 = 0;
                int t2 = 0;
                for (; t2 < t && i + t + f + t2 < nops && is_plain(ops[i + t + f + t2], BFLY_GATE_H); ++t2)
                    T2 |= 1ull << ops[i + t + f + t2].target;
                if (t2 == t && T2 == T) {
                    bfly_op d{};
                    d.kind = BFLY_OP_DIFFUSE;
                    d.aux = T;
                    d.arg_i = 0;
                    out.push_back(d);
                    i += 2 * t + f;
                    continue;
                }
            }
        }
        const int f = match_flip(ops, nops, i, n, mask, value);
        if (f > 0) {
            bfly_op p{};
            p.kind = BFLY_OP_PHASE_FLIP;
            p.aux = mask;
            p.controls = value;
            out.push_back(p);
            i += f;
            continue;
        }
        out.push_back(ops[i]);
        ++i;
    }
    return out;
}

template <typename MayCarry>
void emit_tile_pass(HostPlan& plan, int n, int k, int S, uint64_t tile_mask, const std::vector<PendingGate>& gates,
                    bool factor, const PhaseCarry& carry_in, PhaseCarry* carry_out, MayCarry&& may_carry) {
    Step st{};
    st.kind = S_TILE;
    TilePass& tp = st.tile;
    tp.n = n;
    tp.k = k;
    tp.slots = S;
    const int G = 1 << S;
    int local_of[64];
    int li = 0, oi = 0;
    for (int b = 0; b < n; ++b) {
        if ((tile_mask >> b) & 1) { local_of[b] = li; tp.qpos[li++] = b; }
        else { local_of[b] = -1; tp.opos[oi++] = b; }
    }
    tp.nouter = oi;
    tp.first_sweep = static_cast<int32_t>(plan.sweeps.size());

    const std::vector<Micro> micro = micro_ops(gates, factor, carry_in, carry_out, may_carry);
    int slots[kSlots] = {0, 0, 0, 0};
    int nslots = 0;
    std::vector<const Micro*> cur;
    auto slot_of = [&](int lb) {
        for (int s = 0; s < nslots; ++s)
            if (slots[s] == lb) return s;
        return -1;
    };
    // Sign flips commute and compose by XOR: a Z-family instruction whose
    // selection depends on slot bits only merges into a preceding one.  In
    // 3-slot sweeps a slot-target flip also carries its member factors
    // (-1 where mm1 selects, else 1) in m[0..7], which the kernel reads with
    // the record's other parameters instead of deriving them.
    auto set_factors = [&](TileInstr& in) {
        if (in.kind != T_NEG || in.slot < 0 || S != 3) return;
        for (int m = 0; m < 8; ++m) in.m[m] = ((in.mm1 >> m) & 1u) ? -1.0 : 1.0;
        fill_floats(in.m, in.mf);
    };
    auto push = [&](TileInstr in, int32_t first) {
        in.guard = in.omask != 0 ? 1 : 0;
        const bool uniform_neg = in.kind == T_NEG && in.slot >= 0 && in.lmask == 0 && in.omask == 0;
        if (uniform_neg && static_cast<int32_t>(plan.instrs.size()) > first) {
            TileInstr& prev = plan.instrs.back();
            if (prev.kind == T_NEG && prev.slot >= 0 && prev.lmask == 0 && prev.omask == 0) {
                prev.mm1 = static_cast<uint16_t>(prev.mm1 ^ in.mm1);
                set_factors(prev);
                return;
            }
        }
        set_factors(in);
        plan.instrs.push_back(in);
    };
    // Negation where (index & mask) == pattern: the pattern's slot bits
    // select members, the others are tested per lane / tile.
    auto flip_instr = [&](uint64_t mask, uint64_t pattern) {
        TileInstr in{};
        in.dlocal = -1;
        in.obitpos = -1;
        // negation where (index & mask) == pattern: the pattern's slot
        // bits select members, the others are tested per lane / tile
        in.kind = T_NEG;
        int tq = -1;  // the bit that plays the target: a slot if possible
        for (int b = 0; b < n && tq < 0; ++b)
            if (((mask >> b) & 1) && local_of[b] >= 0 && slot_of(local_of[b]) >= 0) tq = b;
        for (int b = 0; b < n && tq < 0; ++b)
            if (((mask >> b) & 1) && local_of[b] >= 0) tq = b;
        for (int b = 0; b < n && tq < 0; ++b)
            if ((mask >> b) & 1) tq = b;
        uint32_t cvals = 0;
        for (int b = 0; b < n; ++b) {
            if (!((mask >> b) & 1) || b == tq) continue;
            const uint64_t v = (pattern >> b) & 1;
            const int lb = local_of[b];
            if (lb < 0) { in.omask |= 1ull << b; in.oval |= v << b; }
            else if (slot_of(lb) >= 0) { in.cslots |= 1u << slot_of(lb); cvals |= static_cast<uint32_t>(v) << slot_of(lb); }
            else { in.lmask |= 1u << lb; in.lval |= static_cast<uint32_t>(v) << lb; }
        }
        const uint32_t tv = static_cast<uint32_t>((pattern >> tq) & 1);
        const int tl = local_of[tq];
        in.slot = tl >= 0 ? slot_of(tl) : -1;
        uint16_t sel = 0;
        for (uint32_t m = 0; m < static_cast<uint32_t>(G); ++m) {
            if ((m & in.cslots) != cvals) continue;
            if (in.slot >= 0 && ((m >> in.slot) & 1u) != tv) continue;
            sel |= static_cast<uint16_t>(1u << m);
        }
        if (in.slot >= 0) {
            in.mm1 = sel;
        } else {
            if (tl < 0) in.obitpos = static_cast<int8_t>(tq);
            else in.dlocal = tl;
            (tv ? in.mm1 : in.mm0) = sel;
        }
        in.cslots = 0;  // folded into the member sets
        return in;
    };
    auto close = [&]() {
        for (int b = 0; nslots < S && b < k; ++b)
            if (slot_of(b) < 0) slots[nslots++] = b;
        TileSweep sw{};
        for (int s = 0; s < kSlots; ++s) sw.s[s] = s < S ? slots[s] : -1;
        const int eb = plan.dtype == BFLY_C64 ? 4 : 8;
        for (uint32_t m = 0; m < static_cast<uint32_t>(G); ++m) {
            uint32_t d = 0;
            for (int a = 0; a < S; ++a) d |= ((m >> a) & 1u) << slots[a];
            sw.doff[m] = lds_swizzle(d, eb) * static_cast<uint32_t>(eb);
        }
        sw.first = static_cast<int32_t>(plan.instrs.size());
        // a phase of a rotation's qubit just before it in the sweep rides on
        // the rotation's instruction (pre-phase; 1 when there is none)
        std::vector<char> used(cur.size(), 0);
        std::vector<TileInstr> rot_instr(cur.size());
        for (size_t ci = 0; ci < cur.size(); ++ci) {
            const Micro* mu = cur[ci];
            if (mu->kind != Micro::ROT) continue;
            TileInstr& in = rot_instr[ci];
            in.dlocal = -1;
            in.obitpos = -1;
            in.kind = T_ROT;
            in.slot = slot_of(local_of[mu->q]);
            in.m[0] = mu->t;
            in.m[1] = mu->form;
            in.m[2] = 1.0;
            in.m[3] = 0.0;
            if (ci > 0 && !used[ci - 1] && cur[ci - 1]->kind == Micro::PH && cur[ci - 1]->q == mu->q) {
                in.m[2] = cur[ci - 1]->z.r;
                in.m[3] = cur[ci - 1]->z.i;
                used[ci - 1] = 1;
            }
            fill_floats(in.m, in.mf);
        }
        for (size_t ci = 0; ci < cur.size(); ++ci) {
            if (used[ci]) continue;
            const Micro* mu = cur[ci];
            if (mu->kind == Micro::ROT) {
                push(rot_instr[ci], sw.first);
                continue;
            }
            TileInstr in{};
            in.dlocal = -1;
            in.obitpos = -1;
            if (mu->kind == Micro::SCALE) {
                in.kind = T_SCALE;
                in.slot = -1;
                in.m[0] = mu->z.r;
                in.m[1] = mu->z.i;
                fill_floats(in.m, in.mf);
                push(in, sw.first);
                continue;
            }
            const int lt = mu->q >= 0 ? local_of[mu->q] : -1;
            in.slot = lt >= 0 ? slot_of(lt) : -1;
            if (mu->kind == Micro::PH) {
                if (in.slot >= 0) {
                    for (uint32_t m = 0; m < static_cast<uint32_t>(G); ++m)
                        if ((m >> in.slot) & 1u) in.mm1 |= static_cast<uint16_t>(1u << m);
                } else {
                    if (lt < 0) in.obitpos = static_cast<int8_t>(mu->q);
                    else in.dlocal = lt;
                    in.mm1 = 0xFFFF;
                }
                if (is_minus_one(mu->z)) {
                    in.kind = T_NEG;
                } else {
                    in.kind = in.slot >= 0 ? T_PH : T_DIAG;
                    in.m[0] = 1.0;
                    in.m[1] = 0.0;
                    in.m[2] = mu->z.r;
                    in.m[3] = mu->z.i;
                    if (in.kind == T_PH) { in.m[0] = mu->z.r; in.m[1] = mu->z.i; in.m[2] = 0.0; in.m[3] = 0.0; }
                }
                fill_floats(in.m, in.mf);
                push(in, sw.first);
                continue;
            }
            const PendingGate* pg = mu->pg;
            const bfly_op& op = *pg->op;
            if (op.kind == BFLY_OP_PHASE_FLIP) {
                push(flip_instr(op.aux, op.controls), sw.first);
                continue;
            }
            if (pg->c.kind == I_NEG1) {
                // a sign gate is the phase flip of its controls and target all 1
                // (symmetric: the flip's own choice of target prefers a slot)
                const uint64_t qm = op.controls | (1ull << op.target);
                push(flip_instr(qm, qm), sw.first);
                continue;
            }
            const int ik = pg->c.kind;
            in.kind = ik == I_NEG1 ? T_NEG : (pg->c.diag ? T_DIAG : T_PAIR);
            pair_matrix(pg->c, in.m);
            fill_floats(in.m, in.mf);
            for (int b = 0; b < n; ++b) {
                if (!((op.controls >> b) & 1)) continue;
                const int lb = local_of[b];
                if (lb < 0) in.omask |= 1ull << b;
                else if (slot_of(lb) >= 0) in.cslots |= 1u << slot_of(lb);
                else in.lmask |= 1u << lb;
            }
            if (pg->c.diag) {
                if (lt < 0) in.obitpos = static_cast<int8_t>(op.target);
                else if (in.slot < 0) in.dlocal = lt;
                for (uint32_t m = 0; m < static_cast<uint32_t>(G); ++m) {
                    if ((m & in.cslots) != in.cslots) continue;
                    if (in.slot < 0 || ((m >> in.slot) & 1u)) in.mm1 |= static_cast<uint16_t>(1u << m);
                    if (in.slot < 0 || !((m >> in.slot) & 1u)) in.mm0 |= static_cast<uint16_t>(1u << m);
                }
                // only I_DIAG touches target-bit-0 members (P / controlled P / Z leave them as is)
                if (ik != I_DIAG) in.mm0 = 0;
            }
            in.lval = in.lmask;
            in.oval = in.omask;
            push(in, sw.first);
        }
        // rotations on distinct slots that follow each other commute: one
        // record applies them all (3-slot sweeps; the 4-slot kernels keep one
        // record per rotation, which holds their register budget)
        if (S == 3) {
            std::vector<TileInstr> packed;
            for (size_t i = static_cast<size_t>(sw.first); i < plan.instrs.size();) {
                const TileInstr& a = plan.instrs[i];
                size_t j = i;
                uint32_t used_slots = 0;
                while (j < plan.instrs.size() && plan.instrs[j].kind == T_ROT &&
                       !((used_slots >> plan.instrs[j].slot) & 1u)) {
                    used_slots |= 1u << plan.instrs[j].slot;
                    ++j;
                }
                if (j - i < 2) {
                    packed.push_back(a);
                    ++i;
                    continue;
                }
                TileInstr r{};
                bool all_form0 = true;
                for (size_t q = i; q < j; ++q) all_form0 = all_form0 && plan.instrs[q].m[1] == 0.0;
                r.kind = used_slots == 7u && all_form0 ? T_ROT3 : T_ROTS;
                r.slot = -1;
                r.dlocal = -1;
                r.obitpos = -1;
                r.cslots = static_cast<uint8_t>(used_slots);
                for (size_t q = i; q < j; ++q) {
                    const TileInstr& x = plan.instrs[q];
                    const int sl = x.slot;
                    r.m[3 * sl] = x.m[1] != 0.0 ? -x.m[0] : x.m[0];
                    if (x.m[1] != 0.0 && x.m[0] == 0.0) r.m[3 * sl] = -0.0;
                    r.m[3 * sl + 1] = x.m[2];
                    r.m[3 * sl + 2] = x.m[3];
                }
                fill_floats(r.m, r.mf, 16);
                packed.push_back(r);
                i = j;
            }
            plan.instrs.resize(static_cast<size_t>(sw.first));
            plan.instrs.insert(plan.instrs.end(), packed.begin(), packed.end());
        }
        sw.ninstr = static_cast<int32_t>(plan.instrs.size()) - sw.first;
        plan.sweeps.push_back(sw);
        cur.clear();
        nslots = 0;
    };
    // Sweep assignment first, emission second.  A diagonal gate (controlled
    // phase, sign gate, phase flip) commutes with everything but the
    // non-diagonal actions on its own qubits, so it may join any sweep from
    // the one holding the last such action (after it) to the current one: it
    // joins the one where most of its qubits are slots (the latest on a tie),
    // which turns per-lane tests into member sets and lets sign flips merge.
    struct SweepBuild {
        int slots[kSlots];
        int nslots;
        std::vector<const Micro*> ops;
    };
    std::vector<SweepBuild> built;
    built.push_back(SweepBuild{{0, 0, 0, 0}, 0, {}});
    std::vector<int> last_nd(static_cast<size_t>(n), 0);  // sweep of the last non-diagonal action per qubit
    auto need_slot = [&](int lt) {
        SweepBuild& b = built.back();
        for (int s2 = 0; s2 < b.nslots; ++s2)
            if (b.slots[s2] == lt) return;
        if (lt < 0) return;
        if (b.nslots == S) built.push_back(SweepBuild{{0, 0, 0, 0}, 0, {}});
        SweepBuild& c = built.back();
        c.slots[c.nslots++] = lt;
    };
    auto is_slot_in = [&](const SweepBuild& b, int q) {
        const int lt = local_of[q];
        if (lt < 0) return false;
        for (int s2 = 0; s2 < b.nslots; ++s2)
            if (b.slots[s2] == lt) return true;
        return false;
    };
    for (size_t i = 0; i < micro.size(); ++i) {
        const Micro& mu = micro[i];
        const bool nondiag = mu.kind == Micro::ROT || (mu.kind == Micro::GATE && !mu.pg->c.diag);
        if (nondiag) {
            need_slot(local_of[mu.q]);
            last_nd[static_cast<size_t>(mu.q)] = static_cast<int>(built.size()) - 1;
        } else if (mu.kind == Micro::PH && local_of[mu.q] >= 0) {
            // a phase anchored before a rotation of its qubit opens the
            // rotation's slot now; otherwise it only takes a free slot
            bool before_rot = false;
            for (size_t j = i + 1; j < micro.size() && !before_rot; ++j) {
                if (blocks_phase(micro[j], mu.q)) before_rot = true;
                else if (micro[j].kind != Micro::PH) break;
            }
            if (before_rot || built.back().nslots < S) need_slot(local_of[mu.q]);
        } else if (mu.kind == Micro::GATE && mu.pg->c.diag) {
            const bfly_op& op = *mu.pg->op;
            const uint64_t qm = op.kind == BFLY_OP_PHASE_FLIP ? op.aux : (op.controls | (1ull << op.target));
            int lo = 0;
            for (int q = 0; q < n; ++q)
                if ((qm >> q) & 1) lo = std::max(lo, last_nd[static_cast<size_t>(q)]);
            int best = static_cast<int>(built.size()) - 1, best_score = -1;
            for (int b2 = static_cast<int>(built.size()) - 1; b2 >= lo; --b2) {
                int score = 0;
                for (int q = 0; q < n; ++q)
                    if (((qm >> q) & 1) && is_slot_in(built[static_cast<size_t>(b2)], q)) ++score;
                if (score > best_score) { best_score = score; best = b2; }
            }
            built[static_cast<size_t>(best)].ops.push_back(&mu);
            continue;
        }
        built.back().ops.push_back(&mu);
    }
    for (const SweepBuild& b : built) {
        if (b.ops.empty() && b.nslots == 0) continue;
        for (int s2 = 0; s2 < kSlots; ++s2) slots[s2] = b.slots[s2];
        nslots = b.nslots;
        cur = b.ops;
        close();
    }
    tp.nsweeps = static_cast<int32_t>(plan.sweeps.size()) - tp.first_sweep;
    st.gates = static_cast<int32_t>(gates.size());
    plan.steps.push_back(st);
    plan.passes += 1;
}

// -- BLOCK mode -------------------------------------------------------------
using Cplx = std::complex<double>;

// U <- G U for a gate on block slot t with slot controls cm (rows with the
// controls set mix pairwise; diagonal gates scale rows).
void fold_into_block(std::vector<Cplx>& U, const Classified& c, int t, uint32_t cm) {
    double g[8];
    pair_matrix(c, g);
    const Cplx a(g[0], g[1]), b(g[2], g[3]), cc(g[4], g[5]), d(g[6], g[7]);
    const Cplx d0(c.m[0], c.m[1]), d1(c.m[2], c.m[3]);
    for (int i = 0; i < kBlockDim; ++i) {
        if ((i >> t) & 1) continue;
        if ((static_cast<uint32_t>(i) & cm) != cm) continue;
        const int i1 = i | (1 << t);
        for (int col = 0; col < kBlockDim; ++col) {
            Cplx& u0 = U[i * kBlockDim + col];
            Cplx& u1 = U[i1 * kBlockDim + col];
            if (c.diag) {
                if (c.kind == I_DIAG) u0 *= d0;
                u1 *= (c.kind == I_NEG1 ? Cplx(-1.0, 0.0) : d1);
            } else {
                const Cplx v0 = u0, v1 = u1;
                u0 = a * v0 + b * v1;
                u1 = cc * v0 + d * v1;
            }
        }
    }
}

// U <- G U for a dense 2^m unitary (udata entries, row-major, index bit i =
// qubit target + i, apply_unitary's convention) acting on block slots
// `slots[0..m)` for the rows whose control slots `cm` are all set.
void fold_unitary_into_block(std::vector<Cplx>& U, const double* g, int m, const int* slots, uint32_t cm) {
    const int dim = 1 << m;
    uint32_t smask = 0;
    for (int i = 0; i < m; ++i) smask |= 1u << slots[i];
    std::vector<Cplx> v(static_cast<size_t>(dim)), w(static_cast<size_t>(dim));
    std::vector<int> rows(static_cast<size_t>(dim));
    for (int base = 0; base < kBlockDim; ++base) {
        if ((static_cast<uint32_t>(base) & smask) || (static_cast<uint32_t>(base) & cm) != cm) continue;
        for (int e = 0; e < dim; ++e) {
            int r = base;
            for (int i = 0; i < m; ++i)
                if ((e >> i) & 1) r |= 1 << slots[i];
            rows[e] = r;
        }
        for (int col = 0; col < kBlockDim; ++col) {
            for (int e = 0; e < dim; ++e) v[e] = U[rows[e] * kBlockDim + col];
            for (int r = 0; r < dim; ++r) {
                Cplx acc(0.0, 0.0);
                for (int c = 0; c < dim; ++c) acc += Cplx(g[2 * (r * dim + c)], g[2 * (r * dim + c) + 1]) * v[c];
                w[r] = acc;
            }
            for (int e = 0; e < dim; ++e) U[rows[e] * kBlockDim + col] = w[e];
        }
    }
}

// Phase angle of a controlled-phase gate diag(1, e^{i theta}): the exact
// angle the circuit gave for P gates (FP64), pi for Z.
double phase_angle(const bfly_op& op, const Classified& c) {
    if (op.kind == BFLY_OP_GATE && op.gate == BFLY_GATE_P) return op.theta;
    if (c.kind == I_NEG1) return M_PI;
    return std::atan2(c.m[3], c.m[2]);
}

// Predicted execution time of a schedule in units of one full HBM pass over
// the state (2 * 2^n * b bytes at 8 TB/s), from measured MI355X pass times:
// a tile pass costs 0.12 per gate it carries, at least 1.6 (30 qubits: 0.5 ms
// per U3-class gate, 6.9 ms floor, complex128); a dense block 2.2 (complex128,
// 9.5 ms) or 2.4 (complex64, 5.2 ms); a bit reversal 2.9; other steps about
// one pass (a flip touches only its pattern).
double step_cost(const Step& st, int32_t dtype) {
    switch (st.kind) {
        case S_TILE: return std::max(1.6, 0.12 * st.gates);
        case S_BLOCK: return dtype == BFLY_C128 ? 2.2 : 2.4;
        case S_BITREV: return 2.9;
        case S_NORM: return 0.6;
        case S_FLIP: return 0.1;
        default: return 1.2;
    }
}

double schedule_cost(const HostPlan& plan) {
    double c = 0.0;
    for (const Step& st : plan.steps) c += step_cost(st, plan.dtype);
    return c;
}

// Returns BFLY_EUNSUPPORTED as soon as the schedule's predicted cost exceeds
// `budget` (AUTO mode compares it with the tile-pass schedule).
int build_block_plan(HostPlan& plan, int n, int32_t dtype, const bfly_op* ops, uint64_t nops,
                     const bfly_exec_opts& opts, std::string& err, double budget = 1e300) {
    const int b = kBlockQubits;
    // complex128 tiles of 2^11 amplitudes (32 KiB of LDS), complex64 2^12
    const int kmax = dtype == BFLY_C128 ? 11 : 12;
    const int K = n < kmax ? n : kmax;
    const uint64_t all = (1ull << n) - 1;
    // the norm check and renormalisation every 8 blocks is the complex64
    // default; complex128 keeps the reference's unnormalised arithmetic
    const int renorm = opts.renorm_every != 0 ? opts.renorm_every : (dtype == BFLY_C64 ? 8 : -1);
    std::vector<uint32_t> remaining(nops);
    for (uint64_t i = 0; i < nops; ++i) remaining[i] = static_cast<uint32_t>(i);
    std::vector<Classified> cls(nops);
    for (uint64_t i = 0; i < nops; ++i)
        if (ops[i].kind == BFLY_OP_GATE) { cls[i] = classify(ops[i]); plan.gates += 1; }
    auto emit_other = [&](const bfly_op& op) {
        Step st{};
        switch (op.kind) {
            case BFLY_OP_BITREV: {
                uint64_t mask = op.aux;
                if (mask == all) mask = 0;
                if (mask != 0 && popc(mask) <= 1) return;
                st.kind = S_BITREV; st.mask = mask; break;
            }
            case BFLY_OP_SWAP:
                if (op.target == op.arg_i) return;
                st.kind = S_SWAP; st.a = op.target; st.b = op.arg_i; break;
            case BFLY_OP_UNITARY:
                if (op.arg_i <= 0) return;
                st.kind = S_UNITARY; st.a = op.target; st.b = op.arg_i; st.controls = op.controls; st.uoff = op.aux; break;
            case BFLY_OP_PHASE_FLIP:
                st.kind = S_FLIP; st.mask = op.aux; st.controls = op.controls; break;
            case BFLY_OP_DIFFUSE:
                st.kind = S_DIFFUSE; st.mask = op.aux; st.a = op.arg_i; break;
            case BFLY_OP_MEASURE:
                if (op.arg_i == 0) return;
                st.kind = S_MEASURE; st.a = op.arg_i; st.uoff = op.aux; st.b = op.gate; st.controls = op.controls;
                plan.measures += 1;
                break;
            default: return;
        }
        plan.steps.push_back(st);
        if (st.kind != S_FLIP) plan.passes += 1;
    };
    // does a non-diagonal gate on one of `need` follow within the next ops?
    constexpr size_t kLookahead = 256;
    auto upcoming_nd = [&](uint64_t need, size_t from) {
        const size_t lim = std::min(remaining.size(), from + 1 + kLookahead);
        for (size_t j = from + 1; j < lim; ++j) {
            const bfly_op& o = ops[remaining[j]];
            if (o.kind == BFLY_OP_GATE && !cls[remaining[j]].diag && ((need >> o.target) & 1)) return true;
        }
        return false;
    };
    int blocks = 0;
    bool renorm_next = false;
    size_t head = 0;
    constexpr size_t kScanWindow = 8192;
    std::vector<uint32_t> rest;
    // dense unitaries whose qubits (range and controls) fit one block are
    // folded into it like gates
    auto block_unitary = [&](const bfly_op& op) {
        if (op.kind != BFLY_OP_UNITARY || op.arg_i < 1 || op.arg_i > b) return false;
        const uint64_t q = (((1ull << op.arg_i) - 1) << op.target) | op.controls;
        return popc(q) <= b;
    };
    while (head < remaining.size()) {
        const bfly_op& front = ops[remaining[head]];
        if (front.kind != BFLY_OP_GATE && !block_unitary(front)) {
            emit_other(front);
            ++head;
            continue;
        }
        uint64_t Q = 0, frozen = 0, nd_done = 0;  // block qubits; no more non-diagonal; had non-diagonal
        int slot_of[64];
        int qubit_of[kBlockQubits];
        int nq = 0;
        for (int q = 0; q < 64; ++q) slot_of[q] = -1;
        std::vector<Cplx> U(kBlockDim * kBlockDim, Cplx(0.0, 0.0));
        for (int i = 0; i < kBlockDim; ++i) U[i * kBlockDim + i] = 1.0;
        std::vector<BlockPhase> pre, post;
        uint64_t d_nd = 0, d_all = 0;
        size_t placed = 0, scanned_after_defer = 0;
        bool deferred_any = false;
        rest.clear();
        size_t idx = head;
        for (; idx < remaining.size(); ++idx) {
            const uint32_t oi = remaining[idx];
            const bfly_op& op = ops[oi];
            if (deferred_any && ++scanned_after_defer > kScanWindow) break;
            if ((d_nd & all) == all) break;
            if (op.kind != BFLY_OP_GATE && !block_unitary(op)) {
                const uint64_t q = op_qubits(op, n);
                d_nd |= q; d_all |= q;
                rest.push_back(oi);
                deferred_any = true;
                continue;
            }
            if (op.kind == BFLY_OP_UNITARY) {
                const uint64_t range = ((1ull << op.arg_i) - 1) << op.target;
                const uint64_t qall = range | op.controls;
                const uint64_t need = qall & ~Q;
                bool uok = !(range & d_all) && !(qall & d_nd) && !(range & frozen) && nq + popc(need) <= b;
                if (uok) {
                    for (const BlockPhase& ph : pre)
                        if ((ph.omask & need) && (ph.slot >= 0 || popc(ph.omask & need) > 1)) uok = false;
                    for (const BlockPhase& ph : post)
                        if (ph.omask & need) uok = false;
                }
                if (!uok) {
                    d_nd |= range; d_all |= qall;
                    rest.push_back(oi);
                    deferred_any = true;
                    continue;
                }
                for (int q = 0; q < n; ++q)
                    if ((need >> q) & 1) { slot_of[q] = nq; qubit_of[nq++] = q; }
                Q |= need;
                for (BlockPhase& ph : pre)
                    if (ph.omask & need) {
                        ph.slot = slot_of[__builtin_ctzll(ph.omask & need)];
                        ph.omask &= ~need;
                    }
                int slots[kBlockQubits];
                for (int i = 0; i < op.arg_i; ++i) slots[i] = slot_of[op.target + i];
                uint32_t cm = 0;
                for (int q = 0; q < n; ++q)
                    if ((op.controls >> q) & 1) cm |= 1u << slot_of[q];
                fold_unitary_into_block(U, plan.udata.data() + 2 * op.aux, op.arg_i, slots, cm);
                nd_done |= range;
                ++placed;
                continue;
            }
            const Classified& c = cls[oi];
            const uint64_t qall = (1ull << op.target) | op.controls;
            const uint64_t qnd = c.diag ? 0 : (1ull << op.target);
            bool ok = !(qnd & d_all) && !(qall & d_nd);
            // A qubit may join the block only if the phase records already
            // placed stay linear in the slots: a pre-phase naming it outside
            // the block moves it to its slot (at most one slot per record);
            // post-phases (applied after the whole block) forbid it.
            auto can_join = [&](uint64_t need) {
                if (nq + popc(need) > b) return false;
                for (const BlockPhase& ph : pre)
                    if ((ph.omask & need) && (ph.slot >= 0 || popc(ph.omask & need) > 1)) return false;
                for (const BlockPhase& ph : post)
                    if (ph.omask & need) return false;
                return true;
            };
            // Outside qubits that phase records name cannot fill an empty
            // slot later (the kernel tests a record's omask on the column
            // index, where the slot bits are zero): a new record must leave
            // enough unnamed qubits for the slots still empty.
            auto named_mask = [&]() {
                uint64_t m = 0;
                for (const BlockPhase& ph : pre) m |= ph.omask;
                for (const BlockPhase& ph : post) m |= ph.omask;
                return m;
            };
            auto filler_room = [&](uint64_t omask) {
                const uint64_t free = all & ~Q & ~(named_mask() | omask);
                return popc(free) >= b - nq;
            };
            auto join = [&](uint64_t need) {
                for (int q = 0; q < n; ++q)
                    if ((need >> q) & 1) { slot_of[q] = nq; qubit_of[nq++] = q; }
                Q |= need;
                for (BlockPhase& ph : pre)
                    if (ph.omask & need) {
                        ph.slot = slot_of[__builtin_ctzll(ph.omask & need)];
                        ph.omask &= ~need;
                    }
            };
            if (ok && !c.diag) {
                const uint64_t need = qall & ~Q;
                if ((qnd & frozen) || !can_join(need)) {
                    ok = false;
                } else {
                    join(need);
                    uint32_t cm = 0;
                    for (int q = 0; q < n; ++q)
                        if ((op.controls >> q) & 1) cm |= 1u << slot_of[q];
                    fold_into_block(U, c, slot_of[op.target], cm);
                    nd_done |= qnd;
                }
            } else if (ok) {
                if ((qall & ~Q) == 0) {
                    uint32_t cm = 0;
                    for (int q = 0; q < n; ++q)
                        if ((op.controls >> q) & 1) cm |= 1u << slot_of[q];
                    fold_into_block(U, c, slot_of[op.target], cm);
                } else if (upcoming_nd(qall & ~Q, idx) && can_join(qall & ~Q)) {
                    // a qubit that a non-diagonal gate needs soon joins now,
                    // so the phase stays in U instead of pinning the qubit
                    // outside the block
                    join(qall & ~Q);
                    uint32_t cm = 0;
                    for (int q = 0; q < n; ++q)
                        if ((op.controls >> q) & 1) cm |= 1u << slot_of[q];
                    fold_into_block(U, c, slot_of[op.target], cm);
                } else if ((c.kind == I_PHASE1 || c.kind == I_NEG1) && popc(qall & Q) <= 1 &&
                           popc(qall & ~Q) <= 1 && filler_room(qall & ~Q)) {
                    const uint64_t a = qall & Q;
                    BlockPhase ph{};
                    ph.omask = qall & ~Q;
                    ph.slot = a ? slot_of[__builtin_ctzll(a)] : -1;
                    ph.theta = phase_angle(op, c);
                    if (!(a & nd_done)) {
                        pre.push_back(ph);
                    } else {
                        post.push_back(ph);
                        frozen |= a;
                    }
                } else if (can_join(qall & ~Q)) {
                    // any other diagonal joins the block's unitary
                    join(qall & ~Q);
                    uint32_t cm = 0;
                    for (int q = 0; q < n; ++q)
                        if ((op.controls >> q) & 1) cm |= 1u << slot_of[q];
                    fold_into_block(U, c, slot_of[op.target], cm);
                } else {
                    ok = false;
                }
            }
            if (!ok) {
                d_nd |= qnd; d_all |= qall;
                rest.push_back(oi);
                deferred_any = true;
                continue;
            }
            ++placed;
        }
        if (placed == 0) {
            // a gate on more qubits than a block holds runs alone
            Step gst{};
            gst.kind = S_GATE;
            gst.gate = make_gate_launch(front, n);
            plan.steps.push_back(gst);
            plan.passes += 1;
            remaining.erase(remaining.begin() + static_cast<long>(head));
            continue;
        }
        // complete the slots with qubits no phase record names, then the
        // tile with the lowest bits
        uint64_t named = 0;
        for (const BlockPhase& ph : pre) named |= ph.omask;
        for (const BlockPhase& ph : post) named |= ph.omask;
        for (int q = 0; nq < b && q < n; ++q)
            if (!(((Q | named) >> q) & 1)) { slot_of[q] = nq; qubit_of[nq++] = q; Q |= 1ull << q; }
        if (nq < b) {
            err = "internal: no free qubit to complete a block";
            return BFLY_EINVAL;
        }
        uint64_t tile = Q;
        for (int q = 0; popc(tile) < K && q < n; ++q) tile |= 1ull << q;
        Step st{};
        st.kind = S_BLOCK;
        TilePass& tp = st.tile;
        tp.n = n;
        tp.k = K;
        tp.slots = kSlots;
        int li = 0, oi2 = 0, local_of[64];
        for (int q = 0; q < n; ++q) {
            if ((tile >> q) & 1) { local_of[q] = li; tp.qpos[li++] = q; }
            else { local_of[q] = -1; tp.opos[oi2++] = q; }
        }
        tp.nouter = oi2;
        BlockPass& bp = st.block;
        for (int j = 0; j < b; ++j) bp.qloc[j] = local_of[qubit_of[j]];
        bp.renorm = renorm_next ? 1 : 0;
        renorm_next = false;
        bp.mat_off = plan.bmat.size();
        if (dtype == BFLY_C64) {
            // the 64 x 64 real form [[Ur, -Ui], [Ui, Ur]] in FP32
            for (int r = 0; r < 2 * kBlockDim; ++r)
                for (int k2 = 0; k2 < 2 * kBlockDim; ++k2) {
                    const Cplx u = U[(r & 31) * kBlockDim + (k2 & 31)];
                    const double v = (r < 32) == (k2 < 32) ? u.real() : (r < 32 ? -u.imag() : u.imag());
                    plan.bmat.push_back(static_cast<float>(v));
                }
        } else {
            // complex128: the 32 x 64 upper half [Ur, -Ui] in FP64, stored as
            // the same 16 KiB (two 32-bit words per double)
            for (int r = 0; r < kBlockDim; ++r)
                for (int k2 = 0; k2 < 2 * kBlockDim; ++k2) {
                    const Cplx u = U[r * kBlockDim + (k2 & 31)];
                    const double v = k2 < 32 ? u.real() : -u.imag();
                    float w[2];
                    std::memcpy(w, &v, sizeof v);
                    plan.bmat.push_back(w[0]);
                    plan.bmat.push_back(w[1]);
                }
        }
        // phase records summed per table row (FP64), then the device tables
        {
            int row_of[64];
            int nrows = 1;
            uint32_t qm = 0;
            for (int j = 0; j < b; ++j) qm |= 1u << bp.qloc[j];
            for (int i = 0; i < K; ++i)
                if (!((qm >> i) & 1)) row_of[tp.qpos[i]] = nrows++;
            for (int i = 0; i < tp.nouter; ++i) row_of[tp.opos[i]] = nrows++;
            const int ncb = K - b, ncol = 1 << ncb;
            bp.rows_off = plan.brows.size();
            plan.brows.resize(plan.brows.size() + 2 * static_cast<size_t>(nrows) * 6, 0.0);
            bp.tab_off = plan.btab.size();
            plan.btab.resize(plan.btab.size() + 2 * static_cast<size_t>(ncol) * 6 * 2 + 2 * static_cast<size_t>(tp.nouter) * 6,
                             0.0);
            for (int w = 0; w < 2; ++w) {
                double* t = plan.brows.data() + bp.rows_off + static_cast<size_t>(w) * nrows * 6;
                int tab = 0;
                for (const BlockPhase& r : w ? post : pre) {
                    const int row = r.omask ? row_of[__builtin_ctzll(r.omask)] : 0;
                    t[row * 6 + r.slot + 1] += r.theta;
                    tab = 1;
                }
                if (w) bp.post_tab = tab;
                else bp.pre_tab = tab;
                double* cf = plan.btab.data() + bp.tab_off + static_cast<size_t>(w) * ncol * 12;
                for (int col = 0; col < ncol; ++col)
                    for (int j = 0; j < 6; ++j) {
                        double a = t[j];
                        for (int i = 0; i < ncb; ++i)
                            if ((col >> i) & 1) a += t[(1 + i) * 6 + j];
                        cf[(col * 6 + j) * 2] = std::cos(a);
                        cf[(col * 6 + j) * 2 + 1] = std::sin(a);
                    }
                double* oa = plan.btab.data() + bp.tab_off + 2 * static_cast<size_t>(ncol) * 12 +
                             static_cast<size_t>(w) * tp.nouter * 6;
                for (int i = 0; i < tp.nouter; ++i)
                    for (int j = 0; j < 6; ++j) oa[i * 6 + j] = t[(1 + ncb + i) * 6 + j];
            }
        }
        plan.steps.push_back(st);
        plan.passes += 1;
        ++blocks;
        if (budget < 1e300 && schedule_cost(plan) > budget) {
            err = "block schedule over budget";
            return BFLY_EUNSUPPORTED;
        }
        if (renorm > 0 && blocks % renorm == 0) {
            Step nst{};
            nst.kind = S_NORM;
            plan.steps.push_back(nst);
            renorm_next = true;
        }
        std::vector<uint32_t> next;
        next.reserve(rest.size() + (remaining.size() - idx));
        next.insert(next.end(), rest.begin(), rest.end());
        next.insert(next.end(), remaining.begin() + static_cast<long>(idx), remaining.end());
        remaining.swap(next);
        head = 0;
    }
    // a trailing check normalises nothing: drop it
    if (!plan.steps.empty() && plan.steps.back().kind == S_NORM) plan.steps.pop_back();
    return BFLY_OK;
}

}  // namespace

int build_plan_steps(HostPlan& plan, int32_t n, int32_t dtype, const bfly_op* ops, uint64_t nops,
                     const double* udata, uint64_t udata_len, const bfly_exec_opts& opts, std::string& err);

namespace {
// Grover chains: a whole-state diffuser followed (optionally through one
// phase flip) by another whole-state diffuser of the same variant absorbs the
// flip into its apply pass and leaves the next diffuser's partial sums
// (identical bits to running the steps one by one, one HBM read less).
void chain_diffusers(HostPlan& plan) {
    const uint64_t all = (1ull << plan.n) - 1;
    auto whole = [&](const Step& st) { return st.kind == S_DIFFUSE && (st.mask == all || st.mask == 0); };
    std::vector<Step> out;
    out.reserve(plan.steps.size());
    bool odd = false;
    for (size_t i = 0; i < plan.steps.size(); ++i) {
        Step st = plan.steps[i];
        if (whole(st)) {
            const bool ready = !out.empty() && out.back().kind == S_DIFFUSE && (out.back().chain & kChainNext);
            if (!ready) odd = false;
            st.chain = (ready ? kChainReady : 0) | (odd ? kChainOdd : 0);
            size_t j = i + 1;
            bool flip = false;
            if (j < plan.steps.size() && plan.steps[j].kind == S_FLIP) { flip = true; ++j; }
            if (j < plan.steps.size() && whole(plan.steps[j]) && plan.steps[j].a == st.a) {
                st.chain |= kChainNext;
                if (flip) {
                    st.chain |= kChainFlip;
                    st.fmask = plan.steps[i + 1].mask;
                    st.fvalue = plan.steps[i + 1].controls;
                }
                out.push_back(st);
                i = j - 1;          // the flip (if any) is absorbed; the next diffuser comes next
                odd = !odd;
                continue;
            }
            out.push_back(st);
            continue;
        }
        out.push_back(st);
    }
    plan.steps.swap(out);
}
}  // namespace

int build_plan(HostPlan& plan, int32_t n, int32_t dtype, const bfly_op* ops, uint64_t nops,
               const double* udata, uint64_t udata_len, const bfly_exec_opts& opts, std::string& err) {
    const int rc0 = build_plan_steps(plan, n, dtype, ops, nops, udata, udata_len, opts, err);
    if (rc0 == BFLY_OK) chain_diffusers(plan);
    return rc0;
}

int build_plan_steps(HostPlan& plan, int32_t n, int32_t dtype, const bfly_op* ops, uint64_t nops,
                     const double* udata, uint64_t udata_len, const bfly_exec_opts& opts, std::string& err) {
    if (opts.mode == BFLY_MODE_AUTO) {
        bfly_exec_opts fo = opts, bo = opts;
        fo.mode = BFLY_MODE_FUSED;
        bo.mode = BFLY_MODE_BLOCK;
        int rc = build_plan_steps(plan, n, dtype, ops, nops, udata, udata_len, fo, err);
        if (rc || n < 9) return rc;
        // Grover sequences already became flips / diffusers, and circuits whose
        // gates are mostly non-diagonal fill blocks with few gates each: in
        // both cases the tile passes win and the block attempt is skipped
        bool grover = false;
        for (const Step& st : plan.steps) grover |= st.kind == S_FLIP || st.kind == S_DIFFUSE;
        uint64_t ngates = 0, nondiag = 0;
        for (uint64_t i = 0; i < nops; ++i)
            if (ops[i].kind == BFLY_OP_GATE) {
                ++ngates;
                nondiag += classify(ops[i]).diag ? 0 : 1;
            }
        if (!grover && 2 * nondiag <= ngates) {
            HostPlan alt;
            alt.n = n;
            alt.dtype = dtype;
            alt.opts = bo;
            if (udata && udata_len) alt.udata.assign(udata, udata + 2 * udata_len);
            std::string err2;
            const double fused_cost = schedule_cost(plan);
            if (build_block_plan(alt, n, dtype, ops, nops, bo, err2, fused_cost) == BFLY_OK &&
                schedule_cost(alt) < fused_cost)
                plan = std::move(alt);
        }
        plan.opts.mode = BFLY_MODE_AUTO;
        return BFLY_OK;
    }
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
    // BLOCK mode below 9 qubits (fewer than 16 columns of 32) runs fused tiles
    const bool fused = (opts.mode == BFLY_MODE_FUSED || opts.mode == BFLY_MODE_BLOCK) && n >= kMinTileBits;
    if (k > n) k = n;
    if (opts.tile_bits <= 0 && k == n && n >= 10) {
        // A state that fits one tile runs every pass in one workgroup.  Long
        // circuits (at least 48 gates per qubit: Grover searches) run faster
        // on tiles two bits narrower than the state, which spread each pass
        // over several workgroups and let Grover iterations become phase
        // flips and diffusers (measured with the autotuner on 10-13 qubit
        // complex64 Grover: 1.4-2.3x).
        uint64_t ngates = 0;
        for (uint64_t i = 0; i < nops; ++i) ngates += ops[i].kind == BFLY_OP_GATE ? 1 : 0;
        if (ngates >= 48ull * static_cast<uint64_t>(n)) k = n - 2;
    }
    int L = opts.low_bits > 0 ? opts.low_bits : default_low_bits(dtype);
    // the tile kernel stages 16-byte vectors: 2 (f64) or 4 (f32) consecutive
    // amplitudes must be consecutive local indices
    L = std::max(L, dtype == BFLY_C64 ? 2 : 1);
    // at least one tile slot stays free for a non-diagonal target, so the
    // first remaining gate always fits and every pass makes progress
    L = std::min(L, k - 1);
    const int max_gates = opts.max_gates_per_pass > 0 ? opts.max_gates_per_pass : (1 << 30);
    // sweep width: 8-amplitude groups (S = 3) or 16 (S = 4); the threads
    // option names the workgroup size 2^(k - S)
    int S = default_sweep_slots(dtype, k);
    // Generated kernels keep 16-amplitude register groups (4 slots) from 2^12
    // tiles: one LDS round trip per four qubits, the pre-phases of four
    // rotations folded into one factor per member, and the next tile's loads
    // in flight during the sweeps (two waves per SIMD).  Measured on the
    // 30-qubit layered circuit: complex128 342 vs 355 ms, complex64 203 vs
    // 224 ms against 3 slots.
    if ((opts.flags & BFLY_FLAG_JIT) && k >= 12) S = 4;
    if (opts.threads > 0) {
        if (k >= kMinSlots3Bits && opts.threads == (1 << (k - 3))) S = 3;
        else if (opts.threads == (k >= 10 ? (1 << (k - 4)) : 64)) S = 4;
        else {
            err = "threads must be 2^(k-3) or 2^(k-4) for tile bits k (2^(k-3) needs k >= 10)";
            return BFLY_EINVAL;
        }
    }
    if (k < kMinSlots3Bits) S = 4;
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
            case BFLY_OP_PHASE_FLIP:
                st.kind = S_FLIP; st.mask = op.aux; st.controls = op.controls; break;
            case BFLY_OP_DIFFUSE:
                st.kind = S_DIFFUSE; st.mask = op.aux; st.a = op.arg_i; break;
            case BFLY_OP_MEASURE:
                if (op.arg_i == 0) return;
                st.kind = S_MEASURE; st.a = op.arg_i; st.uoff = op.aux; st.b = op.gate; st.controls = op.controls;
                plan.measures += 1;
                break;
            default: return;
        }
        plan.steps.push_back(st);
        if (st.kind != S_FLIP) plan.passes += 1;  // a phase flip touches 2^(n - |mask|) amplitudes
    };

    if (opts.mode == BFLY_MODE_BLOCK && n >= 9) return build_block_plan(plan, n, dtype, ops, nops, opts, err);
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

    // Fused: Grover sequences become phase flips / diffusers (diffusers only
    // when the state spans several tiles: a one-tile circuit keeps every gate
    // in LDS), then greedy commutation-aware pass building.
    for (uint64_t i = 0; i < nops; ++i)
        if (ops[i].kind == BFLY_OP_GATE) plan.gates += 1;
    const bool generic = (opts.flags & BFLY_FLAG_GENERIC_PAIRS) != 0;
    std::vector<bfly_op> as_gates, rewritten;
    bool small_unitaries = false;
    if (udata) as_gates = unitaries_as_gates(ops, nops, udata, small_unitaries);
    if (small_unitaries) {   // plan.gates keeps counting the input's gate ops
        ops = as_gates.data();
        nops = as_gates.size();
    }
    std::vector<bfly_op> merged;
    if (!generic) {
        rewritten = rewrite_grover(ops, nops, n, n > k);
        ops = rewritten.data();
        nops = rewritten.size();
        bool any_merged = false;
        merged = merge_gate_runs(ops, nops, any_merged);
        if (any_merged) {
            ops = merged.data();
            nops = merged.size();
        }
    }
    std::vector<uint32_t> remaining(nops);
    for (uint64_t i = 0; i < nops; ++i) remaining[i] = static_cast<uint32_t>(i);
    std::vector<Classified> cls(nops);
    for (uint64_t i = 0; i < nops; ++i) {
        if (ops[i].kind == BFLY_OP_GATE) cls[i] = classify(ops[i]);
        else if (ops[i].kind == BFLY_OP_PHASE_FLIP) { cls[i] = Classified{}; cls[i].kind = I_NEG1; cls[i].diag = true; }
    }
    auto placeable = [](const bfly_op& op) { return op.kind == BFLY_OP_GATE || op.kind == BFLY_OP_PHASE_FLIP; };
    size_t head = 0;  // remaining[head..] are still to schedule (compacted each pass)
    std::vector<uint32_t> rest;
    std::vector<PendingGate> placed;
    PhaseCarry carry;  // phases waiting for a later tile pass
    constexpr size_t kScanWindow = 8192;  // gates examined past the first deferral
    // per op: qubits acted on non-diagonally (a stand-alone op: all of its
    // qubits) and all qubits it touches; stand-alone ops defer everything
    // on their qubits
    std::vector<uint64_t> op_nd(nops), op_all(nops);
    std::vector<uint8_t> op_place(nops);
    for (uint64_t i = 0; i < nops; ++i) {
        const bfly_op& op = ops[i];
        op_place[i] = placeable(op) ? 1 : 0;
        if (!op_place[i]) {
            op_nd[i] = op_all[i] = op_qubits(op, n);
        } else {
            op_all[i] = op.kind == BFLY_OP_PHASE_FLIP ? op.aux : (1ull << op.target) | op.controls;
            op_nd[i] = cls[i].diag ? 0 : (1ull << op.target);
        }
    }
    // One scan of the ops rem[from..] in program order: an op joins the pass
    // when it commutes with every op deferred before it and its non-diagonal
    // target is a tile qubit.  With `grow` a target outside the tile joins
    // it while the tile has fewer than k qubits (first fit); otherwise the
    // tile is fixed.  Returns the number of ops placed and the final tile;
    // fills the placed ops, the deferred ops and the scan end when asked.
    auto scan_in = [&](const std::vector<uint32_t>& rem, size_t from, uint64_t tile, bool grow,
                       std::vector<PendingGate>* placed_out, std::vector<uint32_t>* rest_out, size_t* end_out,
                       uint64_t* tile_out) -> size_t {
        uint64_t d_nd = 0, d_all = 0;  // qubits of deferred gates: non-diagonal targets / all qubits
        size_t nplaced = 0;
        if (placed_out) placed_out->clear();
        if (rest_out) rest_out->clear();
        size_t scanned_after_defer = 0;
        bool deferred_any = false;
        size_t idx = from;
        for (; idx < rem.size(); ++idx) {
            const uint32_t oi = rem[idx];
            if (deferred_any && ++scanned_after_defer > kScanWindow) break;
            if ((d_nd & all) == all) break;  // every qubit blocked: nothing more can move
            const uint64_t qnd = op_nd[oi], qall = op_all[oi];
            bool ok = op_place[oi] && !(qnd & d_all) && !(qall & d_nd) && static_cast<int>(nplaced) < max_gates;
            if (ok && qnd && !(qnd & tile)) {
                if (grow && popc(tile) < k) tile |= qnd;
                else ok = false;
            }
            if (!ok) {
                d_nd |= qnd; d_all |= qall;
                if (rest_out) rest_out->push_back(oi);
                deferred_any = true;
                continue;
            }
            ++nplaced;
            if (placed_out) placed_out->push_back(PendingGate{&ops[oi], cls[oi]});
        }
        if (end_out) *end_out = idx;
        if (tile_out) *tile_out = tile;
        return nplaced;
    };
    size_t scan_end = head;
    auto scan = [&](uint64_t tile, bool grow, bool record, uint64_t* tile_out) -> size_t {
        return scan_in(remaining, head, tile, grow, record ? &placed : nullptr, record ? &rest : nullptr,
                       record ? &scan_end : nullptr, tile_out);
    };
    // Tile qubits of a pass.  First fit takes the targets in program order.
    // The search (default) also builds the set one qubit at a time, each time
    // adding the qubit that lets the most ops into the pass, keeps the better
    // of the two, then swaps single qubits while that places more ops.  The
    // reference's grid strategies fix the row-local qubits to the low bits
    // (executors.mojo:213-228,949-955); the set here is free, so choosing it
    // well cuts the number of HBM passes (30-qubit layered circuit, 2^12
    // tiles: 55 passes first fit, 38 after the greedy build, 30 after the
    // swaps).
    const bool search = !(opts.flags & BFLY_FLAG_FIRST_FIT) && k < n;
    constexpr int kMaxSwaps = 64;   // bounds the planner's time on long circuits
    auto refine = [&](const std::vector<uint32_t>& rem, size_t from, uint64_t t, size_t cur) {
        for (int swaps = 0; swaps < kMaxSwaps; ++swaps) {
            bool improved = false;
            for (int a = 0; a < n && !improved; ++a) {
                if (!((t >> a) & 1) || ((low_mask >> a) & 1)) continue;
                for (int b = 0; b < n && !improved; ++b) {
                    if ((t >> b) & 1) continue;
                    const uint64_t t2 = (t & ~(1ull << a)) | (1ull << b);
                    const size_t c = scan_in(rem, from, t2, false, nullptr, nullptr, nullptr, nullptr);
                    if (c > cur) { cur = c; t = t2; improved = true; }
                }
            }
            if (!improved) break;
        }
        return t;
    };
    // The search's choice; with `alts`, also the first-fit and greedy sets,
    // each improved by swaps (the candidates of the beam below).
    auto choose_tile_in = [&](const std::vector<uint32_t>& rem, size_t from, std::vector<uint64_t>* alts) {
        uint64_t t_ff = 0;
        scan_in(rem, from, low_mask, true, nullptr, nullptr, nullptr, &t_ff);
        for (int b = 0; popc(t_ff) < k && b < n; ++b) t_ff |= 1ull << b;
        const size_t c_ff = scan_in(rem, from, t_ff, false, nullptr, nullptr, nullptr, nullptr);
        uint64_t t_gr = low_mask;
        size_t c_gr = 0;
        while (popc(t_gr) < k) {
            int bq = -1;
            size_t bc = 0;
            for (int q = 0; q < n; ++q) {
                if ((t_gr >> q) & 1) continue;
                const size_t c = scan_in(rem, from, t_gr | (1ull << q), false, nullptr, nullptr, nullptr, nullptr);
                if (bq < 0 || c > bc) { bq = q; bc = c; }
            }
            t_gr |= 1ull << bq;
            c_gr = bc;
        }
        const uint64_t best = refine(rem, from, c_gr > c_ff ? t_gr : t_ff, std::max(c_gr, c_ff));
        if (alts) {
            alts->clear();
            auto add = [&](uint64_t t) {
                if (std::find(alts->begin(), alts->end(), t) == alts->end()) alts->push_back(t);
            };
            add(best);
            add(t_ff);
            add(t_gr);
            add(refine(rem, from, t_ff, c_ff));
            add(refine(rem, from, t_gr, c_gr));
        }
        return best;
    };
    auto choose_tile = [&]() -> uint64_t { return choose_tile_in(remaining, head, nullptr); };
    // Beam over whole schedules (BFLY_PLAN_BEAM=1): the kBeam partial
    // schedules with the fewest ops left are extended by each candidate set
    // of their next pass, and the first schedule to place every op gives the
    // tile sets of all passes (kept only when it needs fewer passes than the
    // greedy schedule).  30-qubit layered circuit, 2^12 tiles: 28 passes
    // against 30 in 0.06 s of planning, but the same 306 ms per circuit on
    // the generated kernels (the searched passes are bound by their
    // arithmetic, which fewer passes only regroup), so it is off by default.
    std::vector<uint64_t> beam_tiles;
    size_t beam_pos = 0;
    constexpr int kBeam = 3;
    constexpr uint64_t kBeamMaxOps = 20000;
    bool beam_on = false;
    if (const char* b = std::getenv("BFLY_PLAN_BEAM")) beam_on = b[0] == '1';
    if (search && beam_on && nops <= kBeamMaxOps) {
        struct Partial {
            std::vector<uint32_t> rem;
            std::vector<uint64_t> tiles;
        };
        auto first_placeable = [&](const std::vector<uint32_t>& rem) {
            size_t f = 0;
            while (f < rem.size() && !op_place[rem[f]]) ++f;
            return f;
        };
        std::vector<uint64_t> alts;
        std::vector<uint32_t> rest_ids;
        // the greedy schedule in the same model: the beam keeps it only when
        // it needs no more passes
        std::vector<uint64_t> greedy_tiles;
        {
            std::vector<uint32_t> rem = remaining;
            for (;;) {
                const size_t from = first_placeable(rem);
                if (from == rem.size()) break;
                const uint64_t t = choose_tile_in(rem, from, nullptr);
                size_t end = from;
                if (scan_in(rem, from, t, false, nullptr, &rest_ids, &end, nullptr) == 0) break;
                greedy_tiles.push_back(t);
                std::vector<uint32_t> nx = rest_ids;
                nx.insert(nx.end(), rem.begin() + static_cast<long>(end), rem.end());
                rem.swap(nx);
            }
        }
        std::vector<Partial> beam(1);
        beam[0].rem = remaining;
        for (;;) {
            const Partial* done = nullptr;
            for (const Partial& ps : beam)
                if (first_placeable(ps.rem) == ps.rem.size()) { done = &ps; break; }
            if (done) {
                beam_tiles = done->tiles;
                break;
            }
            if (beam[0].tiles.size() >= greedy_tiles.size()) break;   // cannot beat the greedy schedule
            std::vector<Partial> next;
            for (const Partial& ps : beam) {
                const size_t from = first_placeable(ps.rem);
                choose_tile_in(ps.rem, from, &alts);
                for (const uint64_t t : alts) {
                    size_t end = from;
                    if (scan_in(ps.rem, from, t, false, nullptr, &rest_ids, &end, nullptr) == 0) continue;
                    Partial nx;
                    nx.rem = rest_ids;
                    nx.rem.insert(nx.rem.end(), ps.rem.begin() + static_cast<long>(end), ps.rem.end());
                    nx.tiles = ps.tiles;
                    nx.tiles.push_back(t);
                    next.push_back(std::move(nx));
                }
            }
            if (next.empty()) break;   // cannot happen: the search's own choice always places an op
            std::stable_sort(next.begin(), next.end(),
                             [](const Partial& a, const Partial& b) { return a.rem.size() < b.rem.size(); });
            beam.clear();
            for (Partial& ps : next) {
                bool dup = false;
                for (const Partial& q : beam) dup = dup || q.rem == ps.rem;
                if (!dup) beam.push_back(std::move(ps));
                if (static_cast<int>(beam.size()) == kBeam) break;
            }
        }
        if (beam_tiles.empty() || beam_tiles.size() > greedy_tiles.size()) beam_tiles = greedy_tiles;
    }
    while (head < remaining.size()) {
        const bfly_op& front = ops[remaining[head]];
        if (!placeable(front)) {
            emit_standalone(front);
            ++head;
            continue;
        }
        uint64_t tile = low_mask;
        if (search) {
            // the beam's set for this pass (the search's own choice when the
            // beam has none left or its set places nothing here)
            tile = beam_pos < beam_tiles.size() ? beam_tiles[beam_pos++] : choose_tile();
            scan(tile, false, true, &tile);
            if (placed.empty()) {
                tile = choose_tile();
                scan(tile, false, true, &tile);
            }
        } else {
            scan(low_mask, true, true, &tile);
        }
        const size_t idx = scan_end;
        if (placed.empty()) {
            err = "internal planner error: a pass placed no gate";
            return BFLY_EINVAL;
        }
        bool flips_only = true;
        for (const PendingGate& pg : placed) flips_only = flips_only && pg.op->kind == BFLY_OP_PHASE_FLIP;
        // compact: deferred ops keep their order, followed by the unscanned tail
        std::vector<uint32_t> next;
        next.reserve(rest.size() + (remaining.size() - idx));
        next.insert(next.end(), rest.begin(), rest.end());
        next.insert(next.end(), remaining.begin() + static_cast<long>(idx), remaining.end());
        // A trailing phase on q may wait for a later pass when the first
        // remaining op that acts on q non-diagonally is a gate with target q:
        // that gate stays ahead of every later non-diagonal action on q, and
        // the phase commutes with everything before it (diagonal gates, phase
        // flips, gates controlled by q).  A measurement, diffuser or other
        // stand-alone op on q first, or none at all, takes the phase now.
        uint64_t carry_ok = 0, decided = 0;
        for (const uint32_t oi : next) {
            const bfly_op& op = ops[oi];
            if (!placeable(op)) {
                decided |= all;   // conservatively every qubit
                break;
            }
            if (op.kind == BFLY_OP_GATE && !cls[oi].diag) {
                const uint64_t t = 1ull << op.target;
                if (!(decided & t)) carry_ok |= t;
                decided |= t;
                if ((decided & all) == all) break;
            }
        }
        if (flips_only) {
            // a pass of phase flips alone: each touches 2^(n - |mask|) amplitudes;
            // the phases waiting for later passes commute with them
            for (const PendingGate& pg : placed) emit_standalone(*pg.op);
        } else {
            // complete the tile with the lowest unused qubits (any k-subset works)
            for (int b = 0; popc(tile) < k && b < n; ++b) tile |= 1ull << b;
            PhaseCarry carry_next;
            emit_tile_pass(plan, n, k, S, tile, placed, !generic, carry, generic ? nullptr : &carry_next,
                           [&](int q) { return ((carry_ok >> q) & 1) != 0; });
            carry.swap(carry_next);
        }
        remaining.swap(next);
        head = 0;
    }
    if (!carry.empty()) {
        err = "internal planner error: a carried phase was not applied";
        return BFLY_EINVAL;
    }
    return BFLY_OK;
}

}  // namespace bfly
