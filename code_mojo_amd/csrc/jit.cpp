// jit.cpp -- straight-line tile-pass kernels generated at plan time.
//
// The interpreting tile kernel (tile_pass_kernel, kernels.hip) reads every
// sweep and instruction record at run time and branches on its kind, slot,
// form and masks for every register group of every tile.  For a plan that
// is executed many times (the same circuit structure with new angles, or a
// benchmark), BFLY_FLAG_JIT instead writes each tile pass as its own kernel:
// slot positions, LDS member offsets, gate kinds, rotation forms, sign
// patterns and lane / tile masks become compile-time constants, and only the
// gate parameters are read at run time, from a compact per-pass array.  The
// arithmetic is the interpreter's own (the helpers of tile_ops.h, called with
// the same parameter values), so both kernels give the same amplitudes.
//
// Sources are compiled with hipRTC for gfx950 when the plan is created (host
// only, several programs in parallel), cached process-wide by source text,
// and loaded as modules on the device at the plan's first execution.  The
// grouping of gates into passes and sweeps is the planner's
// (execute_grid_fused's row-local groups generalised, executors.mojo:1062-1093).
#include <hip/hip_runtime_api.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bfly_plan.h"
#include "jit.h"

namespace bfly {

struct JitCode {
    std::vector<char> code;
};

namespace {

// The text of tile_ops.h, embedded at build time.
const char kTileOps[] =
#include "tile_ops_src.inc"
    ;

constexpr int tile_threads_h(int k, int s) { return k - s >= 6 ? (1 << (k - s)) : 64; }
constexpr int tile_waves_per_eu_h(int elem, int k, int s) {
    const int lds = 2 * elem << k;
    const int blocks = lds > 80 * 1024 ? 1 : 2;
    const int w = blocks * (tile_threads_h(k, s) / 64) / 4;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

struct Out {
    std::string s;
    template <typename... A>
    void operator()(const char* fmt, A... a) {
        const int need = std::snprintf(nullptr, 0, fmt, a...);
        if (need <= 0) return;
        const size_t at = s.size();
        s.resize(at + static_cast<size_t>(need) + 1);
        std::snprintf(&s[at], static_cast<size_t>(need) + 1, fmt, a...);
        s.resize(at + static_cast<size_t>(need));
    }
};

// Expression depositing the bits [lo, hi) of `var` at global positions pos[i]
// (ascending runs become one shift and mask each).
std::string deposit_expr(const char* var, const int32_t* pos, int lo, int hi) {
    std::string e;
    char b[160];
    int i = lo;
    while (i < hi) {
        int j = i + 1;
        while (j < hi && pos[j] == pos[j - 1] + 1) ++j;
        const int len = j - i;
        const unsigned long long mask = (len >= 64) ? ~0ull : ((1ull << len) - 1);
        std::snprintf(b, sizeof b, "%s(((((unsigned long long)(%s)) >> %d) & 0x%llxull) << %d)", e.empty() ? "" : " | ",
                      var, i, mask, pos[i]);
        e += b;
        i = j;
    }
    return e.empty() ? std::string("0ull") : e;
}

unsigned long long deposit_const(unsigned long long v, const int32_t* pos, int nbits) {
    unsigned long long g = 0;
    for (int i = 0; i < nbits; ++i) g |= ((v >> i) & 1ull) << pos[i];
    return g;
}


// -- per-sweep IR -----------------------------------------------------------
// A sweep's records become ROT (one rotation on a slot), PHASE (slot bit-1
// members times z), NEG (exact sign flips of a member set under a lane / tile
// condition) and TEXT (other forms, emitted as written).  Sign flips whose
// members all have slot a's bit set fold into an adjacent rotation or phase
// on slot a: a flip of the pair's v member before a rotation is z -> -z, after a
// form-0 rotation (t, z) -> (-t, -z), on a phase z -> -z.  Negation is exact,
// so the amplitudes equal those of the separate flip.
struct Flip {
    std::string cond;    // "" = always
    unsigned members;    // flipped members (v members of the target slot)
};

struct IrOp {
    enum Kind { ROT, PHASE, NEG, SCALE, TEXT } kind = TEXT;
    int slot = -1;
    int ct = -1, cr = -1, ci = -1;
    double zr = 1.0, zi = 0.0;     // ROT: pre-phase, PHASE: phase, SCALE: factor (FP64 values)
    bool hoisted = false;          // ROT: pre-phase applied by the sweep's member factors
    bool form1 = false;
    unsigned members = 0;          // NEG
    std::string cond;              // NEG: condition ("" = always)
    std::vector<Flip> pre, post;   // ROT / PHASE: folded flips
    unsigned nd = 0;               // TEXT: slots acted on non-diagonally
    std::string text;
};

std::string and_cond(const std::string& a, const std::string& b) {
    if (a.empty()) return b;
    if (b.empty()) return a;
    return "(" + a + " && " + b + ")";
}

template <typename Cst>
void lower_instr(std::vector<IrOp>& ir, const TileInstr& in, int S, bool f32, Cst&& cst) {
    const int G = 1 << S;
    const double* md = in.m;
    const float* mf = in.mf;
    auto C = [&](int i) { return cst(md[i], mf[i]); };
    char b[256];
    std::string cond;
    if (in.guard) {
        std::snprintf(b, sizeof b, "((base & 0x%llxull) == 0x%llxull)", static_cast<unsigned long long>(in.omask),
                      static_cast<unsigned long long>(in.oval));
        cond = b;
    }
    if (in.lmask) {
        std::snprintf(b, sizeof b, "((jb & %uu) == %uu)", in.lmask, in.lval);
        cond = and_cond(cond, b);
    }
    auto rot = [&](int slot, int ct, int cr, int ci, bool form1, double zr, double zi) {
        IrOp op;
        op.kind = IrOp::ROT;
        op.slot = slot;
        op.ct = ct;
        op.cr = cr;
        op.ci = ci;
        op.zr = zr;
        op.zi = zi;
        op.form1 = form1;
        if (!cond.empty()) {
            // a conditional rotation stays as written
            op.kind = IrOp::TEXT;
            op.nd = 1u << slot;
            std::snprintf(b, sizeof b, "    if %s slot_rot<T, %d, %d>(ar, ai, C[%d], %s, C[%d], C[%d]);\n",
                          cond.c_str(), G, slot, ct, form1 ? "true" : "false", cr, ci);
            op.text = b;
        }
        ir.push_back(op);
    };
    auto neg = [&](unsigned members, const std::string& c) {
        if (!members) return;
        IrOp op;
        op.kind = IrOp::NEG;
        op.members = members;
        op.cond = c;
        ir.push_back(op);
    };
    auto text = [&](const std::string& t, unsigned nd) {
        IrOp op;
        op.kind = IrOp::TEXT;
        op.nd = nd;
        op.text = cond.empty() ? t : "    if " + cond + " {\n" + t + "    }\n";
        ir.push_back(op);
    };
    const int kind = in.kind;
    if (S == 3 && kind == T_ROT3) {
        for (int a = 0; a < 3; ++a) {
            const int ct = C(3 * a), cr = C(3 * a + 1), ci = C(3 * a + 2);
            rot(a, ct, cr, ci, false, md[3 * a + 1], md[3 * a + 2]);
        }
    } else if (kind == T_ROT) {
        const bool form1 = f32 ? mf[1] != 0.0f : md[1] != 0.0;
        const int ct = C(0), cr = C(2), ci = C(3);
        rot(in.slot, ct, cr, ci, form1, md[2], md[3]);
    } else if (S == 3 && kind == T_ROTS) {
        for (int a = 0; a < S; ++a) {
            if (!((in.cslots >> a) & 1u)) continue;
            const double hd = md[3 * a];
            const float hf = mf[3 * a];
            const bool form1 = f32 ? std::signbit(hf) : std::signbit(hd);
            const int ct = cst(std::fabs(hd), std::fabs(hf));
            const int cr = C(3 * a + 1), ci = C(3 * a + 2);
            rot(a, ct, cr, ci, form1, md[3 * a + 1], md[3 * a + 2]);
        }
    } else if (kind == T_PH) {
        const int cr = C(0), ci = C(1);
        if (cond.empty()) {
            IrOp op;
            op.kind = IrOp::PHASE;
            op.slot = in.slot;
            op.cr = cr;
            op.ci = ci;
            op.zr = md[0];
            op.zi = md[1];
            ir.push_back(op);
        } else {
            std::snprintf(b, sizeof b, "    slot_phase<T, %d, %d>(ar, ai, C[%d], C[%d]);\n", G, int(in.slot), cr, ci);
            text(b, 0);
        }
    } else if (kind == T_SCALE && cond.empty()) {
        // a scalar: the sweep folds it into its member factors
        IrOp op;
        op.kind = IrOp::SCALE;
        op.zr = md[0];
        op.zi = md[1];
        ir.push_back(op);
    } else if (kind == T_SCALE) {
        const int cr = C(0), ci = C(1);
        const bool real = f32 ? mf[1] == 0.0f : md[1] == 0.0;
        if (real)
            std::snprintf(b, sizeof b,
                          "    { const T z = C[%d];\n#pragma unroll\n      for (int m = 0; m < %d; ++m) { ar[m] *= z; ai[m] *= z; } }\n",
                          cr, G);
        else
            std::snprintf(b, sizeof b, "    scale_all<T, %d>(ar, ai, C[%d], C[%d]);\n", G, cr, ci);
        text(b, 0);
    } else if (kind == T_PAIR) {
        int cg[8];
        for (int i = 0; i < 8; ++i) cg[i] = C(i);
        std::string t;
        std::snprintf(b, sizeof b, "    { const T g[8] = {C[%d], C[%d], C[%d], C[%d], C[%d], C[%d], C[%d], C[%d]};\n",
                      cg[0], cg[1], cg[2], cg[3], cg[4], cg[5], cg[6], cg[7]);
        t += b;
        // pairs whose slot controls do not hold keep their values
        const int sl = in.slot;
        for (int p = 0; p < G / 2; ++p) {
            const int lo_mask = (1 << sl) - 1;
            const int m0 = ((p & ~lo_mask) << 1) | (p & lo_mask);
            const int m1 = m0 | (1 << sl);
            if ((static_cast<unsigned>(m0) & in.cslots) != in.cslots) continue;
            std::snprintf(b, sizeof b, "      pair_gen(ar[%d], ai[%d], ar[%d], ai[%d], g);\n", m0, m0, m1, m1);
            t += b;
        }
        t += "    }\n";
        text(t, 1u << sl);
    } else {
        // T_NEG / T_DIAG: member selections sel1 (target bit 1) / sel0
        const unsigned sel1 = in.mm1, sel0 = in.mm0;
        if (in.slot < 0) {
            std::string bit1;
            if (in.dlocal >= 0) std::snprintf(b, sizeof b, "(((jb >> %d) & 1u) != 0u)", int(in.dlocal));
            else std::snprintf(b, sizeof b, "(((base >> %d) & 1ull) != 0ull)", int(in.obitpos));
            bit1 = b;
            if (kind == T_NEG) {
                // member factor F (= -1 where the bit is 1) for mm1 only, -F for mm0 only, -1 for both
                unsigned both = 0, only1 = 0, only0 = 0;
                for (int m = 0; m < G; ++m) {
                    const unsigned f1 = (sel1 >> m) & 1u, f0 = (sel0 >> m) & 1u;
                    if (f1 && f0) both |= 1u << m;
                    else if (f1) only1 |= 1u << m;
                    else if (f0) only0 |= 1u << m;
                }
                neg(both, cond);
                neg(only1, and_cond(cond, bit1));
                neg(only0, and_cond(cond, "!" + bit1));
            } else {
                const int c0r = C(0), c0i = C(1), c1r = C(2), c1i = C(3);
                std::string t = "    { const bool bit = " + bit1 + ";\n";
                for (int m = 0; m < G; ++m) {
                    const unsigned s1 = (sel1 >> m) & 1u, s0 = (sel0 >> m) & 1u;
                    if (!s1 && !s0) continue;
                    std::snprintf(b, sizeof b,
                                  "      if (bit ? %u : %u) cmul(ar[%d], ai[%d], bit ? C[%d] : C[%d], bit ? C[%d] : C[%d]);\n",
                                  s1, s0, m, m, c1r, c0r, c1i, c0i);
                    t += b;
                }
                t += "    }\n";
                text(t, 0);
            }
        } else if (kind == T_NEG) {
            // exact sign flips: S = 3 records carry the member factors, S = 4 the selection
            unsigned members = 0;
            for (int m = 0; m < G; ++m) {
                const bool ng = S == 3 ? (f32 ? mf[m] < 0.0f : md[m] < 0.0) : ((sel1 >> m) & 1u) != 0;
                if (ng) members |= 1u << m;
            }
            neg(members, cond);
        } else {
            const int c0r = C(0), c0i = C(1), c1r = C(2), c1i = C(3);
            std::string t;
            for (int m = 0; m < G; ++m) {
                if ((sel1 >> m) & 1u) std::snprintf(b, sizeof b, "    cmul(ar[%d], ai[%d], C[%d], C[%d]);\n", m, m, c1r, c1i);
                else if ((sel0 >> m) & 1u) std::snprintf(b, sizeof b, "    cmul(ar[%d], ai[%d], C[%d], C[%d]);\n", m, m, c0r, c0i);
                else continue;
                t += b;
            }
            text(t, 0);
        }
    }
}

// Member set M is unchanged by flipping slot bit c.
bool independent_of(unsigned M, int c, int G) {
    for (int m = 0; m < G; ++m)
        if (((M >> m) & 1u) != ((M >> (m ^ (1 << c))) & 1u)) return false;
    return true;
}

// A sign flip of members M commutes with op x (the sweep's ops act on slots
// only; lane and tile conditions are constant within a sweep).
bool neg_commutes(unsigned M, const IrOp& x, int G) {
    if (x.kind == IrOp::ROT) return independent_of(M, x.slot, G);
    if (x.kind == IrOp::TEXT) {
        for (int c = 0; (1 << c) < G; ++c)
            if (((x.nd >> c) & 1u) && !independent_of(M, c, G)) return false;
    }
    return true;
}

void fold_signs(std::vector<IrOp>& ir, int S) {
    const int G = 1 << S;
    for (size_t k = 0; k < ir.size(); ++k) {
        if (ir[k].kind != IrOp::NEG) continue;
        const unsigned M = ir[k].members;
        bool done = false;
        for (int a = 0; a < S && !done; ++a) {
            unsigned bit_a = 0;
            for (int m = 0; m < G; ++m)
                if ((m >> a) & 1) bit_a |= 1u << m;
            if (M & ~bit_a) continue;   // a u member flips: not a v-member phase of slot a
            // backward: the flip happened after the op
            for (size_t j = k; j-- > 0;) {
                IrOp& x = ir[j];
                if ((x.kind == IrOp::ROT || x.kind == IrOp::PHASE) && x.slot == a) {
                    if (x.kind == IrOp::ROT && x.form1) break;
                    x.post.push_back(Flip{ir[k].cond, M});
                    done = true;
                    break;
                }
                if (!neg_commutes(M, x, G)) break;
            }
            if (done) break;
            // forward: the flip happens before the op
            for (size_t j = k + 1; j < ir.size(); ++j) {
                IrOp& x = ir[j];
                if ((x.kind == IrOp::ROT || x.kind == IrOp::PHASE) && x.slot == a) {
                    x.pre.push_back(Flip{ir[k].cond, M});
                    done = true;
                    break;
                }
                if (!neg_commutes(M, x, G)) break;
            }
        }
        if (done) {
            ir.erase(ir.begin() + static_cast<std::ptrdiff_t>(k));
            --k;
        }
    }
}

// XOR of the conditions of the flips that contain member m, as an expression
// ("" when none applies, "true" when an unconditional flip is odd).
std::string flip_expr(const std::vector<Flip>& fl, int m) {
    std::vector<std::string> cs;
    bool always = false;
    for (const Flip& f : fl) {
        if (!((f.members >> m) & 1u)) continue;
        if (f.cond.empty()) always = !always;
        else cs.push_back(f.cond);
    }
    std::string e;
    for (const std::string& c : cs) e = e.empty() ? c : "(" + e + " != " + c + ")";
    if (always) e = e.empty() ? std::string("true") : "!" + e;
    return e;
}

// Member factors: every diagonal action of a sweep that commutes with all
// the ops still ahead of it (pre-phases of rotations on slots nothing earlier
// acts on non-diagonally, phases, unconditional sign flips, scalars) is
// applied first, as one complex factor per register member, computed in FP64
// on the host.  A three-rotation sweep then costs 7 complex products and
// 12 four-FMA pair rotations instead of 12 pre-phased pair rotations (about
// a fifth fewer FP64 instructions).  Returns the per-member factors; ops
// folded into them are marked (ROT: pre-phase and pre flips) or dropped.
struct Cxd {
    double r, i;
};
Cxd cxmul(Cxd a, Cxd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

bool flips_unconditional(const std::vector<Flip>& fl) {
    for (const Flip& f : fl)
        if (!f.cond.empty()) return false;
    return true;
}

std::vector<Cxd> hoist_diagonals(std::vector<IrOp>& ir, int S) {
    const int G = 1 << S;
    std::vector<Cxd> F(static_cast<size_t>(G), Cxd{1.0, 0.0});
    // BFLY_JIT_HOIST=0 keeps every diagonal action in place (measurement)
    if (const char* h = std::getenv("BFLY_JIT_HOIST"))
        if (h[0] == '0') return F;
    std::vector<IrOp> out;
    unsigned nd_before = 0;   // slots acted on non-diagonally by the ops kept so far
    auto commutes_with_kept = [&](unsigned M) {
        for (const IrOp& y : out)
            if (!neg_commutes(M, y, G)) return false;
        return true;
    };
    auto flips_commute = [&](const std::vector<Flip>& fl) {
        for (const Flip& f : fl)
            if (!commutes_with_kept(f.members)) return false;
        return true;
    };
    auto apply_flips = [&](const std::vector<Flip>& fl) {
        for (const Flip& f : fl)
            for (int m = 0; m < G; ++m)
                if ((f.members >> m) & 1u) F[static_cast<size_t>(m)] = Cxd{-F[static_cast<size_t>(m)].r, -F[static_cast<size_t>(m)].i};
    };
    auto apply_phase = [&](int slot, double zr, double zi) {
        for (int m = 0; m < G; ++m)
            if ((m >> slot) & 1) F[static_cast<size_t>(m)] = cxmul(F[static_cast<size_t>(m)], Cxd{zr, zi});
    };
    for (IrOp& x : ir) {
        const unsigned sb = x.slot >= 0 ? 1u << x.slot : 0u;
        if (x.kind == IrOp::ROT) {
            if (!(nd_before & sb) && flips_unconditional(x.pre) && flips_commute(x.pre)) {
                apply_phase(x.slot, x.zr, x.zi);
                apply_flips(x.pre);
                x.pre.clear();
                x.hoisted = true;
            }
            nd_before |= sb;
            out.push_back(x);
        } else if (x.kind == IrOp::PHASE) {
            if (!(nd_before & sb) && flips_unconditional(x.pre) && flips_unconditional(x.post) &&
                flips_commute(x.pre) && flips_commute(x.post)) {
                apply_flips(x.pre);
                apply_phase(x.slot, x.zr, x.zi);
                apply_flips(x.post);
                continue;
            }
            out.push_back(x);
        } else if (x.kind == IrOp::NEG) {
            if (x.cond.empty() && commutes_with_kept(x.members)) {
                apply_flips({Flip{"", x.members}});
                continue;
            }
            out.push_back(x);
        } else if (x.kind == IrOp::SCALE) {
            for (Cxd& f : F) f = cxmul(f, Cxd{x.zr, x.zi});
            continue;
        } else {
            nd_before |= x.nd;
            out.push_back(x);
        }
    }
    ir.swap(out);
    return F;
}

template <typename Cst>
void emit_ops(Out& o, std::vector<IrOp>& ir, int S, Cst&& cst) {
    const int G = 1 << S;
    int var = 0;
    const std::vector<Cxd> F = hoist_diagonals(ir, S);
    for (int m = 0; m < G; ++m) {
        const Cxd f = F[static_cast<size_t>(m)];
        if (f.r == 1.0 && f.i == 0.0) continue;
        if (f.r == -1.0 && f.i == 0.0) {
            o("    ar[%d] = -ar[%d]; ai[%d] = -ai[%d];\n", m, m, m, m);
        } else if (f.i == 0.0) {
            const int c = cst(f.r, static_cast<float>(f.r));
            o("    ar[%d] *= C[%d]; ai[%d] *= C[%d];\n", m, c, m, c);
        } else {
            const int cr = cst(f.r, static_cast<float>(f.r)), ci = cst(f.i, static_cast<float>(f.i));
            o("    cmul(ar[%d], ai[%d], C[%d], C[%d]);\n", m, m, cr, ci);
        }
    }
    for (const IrOp& x : ir) {
        if (x.kind == IrOp::ROT && x.hoisted) {
            // pre-phase already applied.  An unconditional post flip of v is
            // z = -1 with t -> -t; a lane / tile conditional one negates v
            // after the rotation (one sign-bit XOR per double).
            const int lo_mask = (1 << x.slot) - 1;
            o("    {\n");
            bool have_pos = false, have_neg = false;
            const int vp = var++, vn = var++;
            std::map<std::string, int> masks;
            for (int p = 0; p < G / 2; ++p) {
                const int m0 = ((p & ~lo_mask) << 1) | (p & lo_mask);
                const int m1 = m0 | (1 << x.slot);
                const std::string fe = flip_expr(x.post, m1);
                const bool flip = fe == "true";
                if (flip && !have_neg) { o("      const T t%d = -C[%d];\n", vn, x.ct); have_neg = true; }
                if (!flip && !have_pos) { o("      const T t%d = C[%d];\n", vp, x.ct); have_pos = true; }
                o("      rot_pair%dz<%d>(ar[%d], ai[%d], ar[%d], ai[%d], t%d);\n", x.form1 ? 1 : 0, flip ? -1 : 1, m0, m0,
                  m1, m1, flip ? vn : vp);
                if (!fe.empty() && !flip) {
                    auto it = masks.find(fe);
                    int mv;
                    if (it == masks.end()) {
                        mv = var++;
                        masks.emplace(fe, mv);
                        o("      const unsigned sg%d = %s ? 0x80000000u : 0u;\n", mv, fe.c_str());
                    } else {
                        mv = it->second;
                    }
                    o("      ar[%d] = xsign(ar[%d], sg%d); ai[%d] = xsign(ai[%d], sg%d);\n", m1, m1, mv, m1, m1, mv);
                }
            }
            o("    }\n");
            continue;
        }
        if (x.kind == IrOp::SCALE) {
            // not folded into the member factors (measurement switch)
            const int cr = cst(x.zr, static_cast<float>(x.zr));
            if (x.zi == 0.0) {
                o("    {\n#pragma unroll\n      for (int m = 0; m < %d; ++m) { ar[m] *= C[%d]; ai[m] *= C[%d]; }\n    }\n", G, cr, cr);
            } else {
                const int ci = cst(x.zi, static_cast<float>(x.zi));
                o("    scale_all<T, %d>(ar, ai, C[%d], C[%d]);\n", G, cr, ci);
            }
            continue;
        }
        if (x.kind == IrOp::TEXT) {
            o("%s", x.text.c_str());
        } else if (x.kind == IrOp::NEG) {
            if (!x.cond.empty()) o("    if %s {\n", x.cond.c_str());
            for (int m = 0; m < G; ++m)
                if ((x.members >> m) & 1u) o("    ar[%d] = -ar[%d]; ai[%d] = -ai[%d];\n", m, m, m, m);
            if (!x.cond.empty()) o("    }\n");
        } else if (x.kind == IrOp::ROT) {
            const int lo_mask = (1 << x.slot) - 1;
            o("    {\n");
            std::map<std::pair<std::string, std::string>, int> vars;   // (z flip, t flip) -> variable set
            for (int p = 0; p < G / 2; ++p) {
                const int m0 = ((p & ~lo_mask) << 1) | (p & lo_mask);
                const int m1 = m0 | (1 << x.slot);
                const std::string fpost = flip_expr(x.post, m1);
                std::string fz = flip_expr(x.pre, m1);
                // z flips with pre XOR post, t with post
                if (!fpost.empty()) fz = fz.empty() ? fpost : "(" + fz + " != " + fpost + ")";
                auto key = std::make_pair(fz, fpost);
                auto it = vars.find(key);
                int v;
                if (it == vars.end()) {
                    v = var++;
                    vars.emplace(key, v);
                    auto sel = [&](const std::string& f, int c) {
                        char b[64];
                        if (f.empty()) std::snprintf(b, sizeof b, "C[%d]", c);
                        else if (f == "true") std::snprintf(b, sizeof b, "-C[%d]", c);
                        else return "(" + f + " ? -C[" + std::to_string(c) + "] : C[" + std::to_string(c) + "])";
                        return std::string(b);
                    };
                    o("      const T t%d = %s, zr%d = %s, zi%d = %s;\n", v, sel(fpost, x.ct).c_str(), v,
                      sel(fz, x.cr).c_str(), v, sel(fz, x.ci).c_str());
                } else {
                    v = it->second;
                }
                o("      rot_pair%d(ar[%d], ai[%d], ar[%d], ai[%d], t%d, zr%d, zi%d);\n", x.form1 ? 1 : 0, m0, m0, m1,
                  m1, v, v, v);
            }
            o("    }\n");
        } else {   // PHASE
            o("    {\n");
            std::map<std::string, int> vars;
            for (int m = 0; m < G; ++m) {
                if (!((m >> x.slot) & 1)) continue;
                std::string f = flip_expr(x.pre, m);
                const std::string f2 = flip_expr(x.post, m);
                if (!f2.empty()) f = f.empty() ? f2 : "(" + f + " != " + f2 + ")";
                auto it = vars.find(f);
                int v;
                if (it == vars.end()) {
                    v = var++;
                    vars.emplace(f, v);
                    if (f.empty()) o("      const T zr%d = C[%d], zi%d = C[%d];\n", v, x.cr, v, x.ci);
                    else if (f == "true") o("      const T zr%d = -C[%d], zi%d = -C[%d];\n", v, x.cr, v, x.ci);
                    else
                        o("      const T zr%d = %s ? -C[%d] : C[%d], zi%d = %s ? -C[%d] : C[%d];\n", v, f.c_str(), x.cr,
                          x.cr, v, f.c_str(), x.ci, x.ci);
                } else {
                    v = it->second;
                }
                o("      cmul(ar[%d], ai[%d], zr%d, zi%d);\n", m, m, v, v);
            }
            o("    }\n");
        }
    }
}

}  // namespace

bool jit_supported(const HostPlan& hp, const Step& st) {
    if (st.kind != S_TILE) return false;
    const TilePass& P = st.tile;
    const size_t elem = hp.dtype == BFLY_C64 ? 4 : 8;
    if (2 * elem * (size_t(1) << P.k) > 128 * 1024) return false;   // LDS of one workgroup (gfx950: 160 KiB)
    return P.slots == 3 || P.slots == 4;
}

// One tile pass as a kernel `bfly_tile_pass(T* re, T* im, const T* C)`.
JitSource jit_generate(const HostPlan& hp, const Step& st) {
    const TilePass& P = st.tile;
    const bool f32 = hp.dtype == BFLY_C64;
    const char* T = f32 ? "float" : "double";
    const int elem = f32 ? 4 : 8;
    const int K = P.k, S = P.slots, G = 1 << S;
    const unsigned TS = 1u << K;
    const int NT = tile_threads_h(K, S);
    const unsigned NG = TS / G;
    const int WPE = tile_waves_per_eu_h(elem, K, S);
    // Split round trips (complex128): the LDS tile holds one component at a
    // time, so a sweep boundary moves re, then im, through a buffer of half
    // the size (three barriers instead of one: re out, re in, im out, im in)
    // and three blocks fit a CU instead of two (the registers allow three
    // waves per SIMD).  Needs one group of members per thread (the barriers
    // sit inside the sweeps).  A standalone probe of the same pass shape (9
    // sweeps of about 250 FP64 VALU per thread, 2^18 tiles): 11.96 ms with
    // 16-byte round trips at two blocks per CU, 9.76 ms split at three, 12.14
    // ms split at two.  Taken in passes of 6 or more sweeps: light passes are
    // bound by their HBM stream, where the work queue's next-tile loads (whose
    // registers do not fit three blocks per CU) pay more.  Measured on the
    // 30-qubit layered circuit, ms per circuit (searched / first-fit tile
    // qubits; one box): split in every pass 270.6 / 408.9, from 6 sweeps
    // 271.5 / 391.8, from 10 sweeps 282.6 / 394.0.  BFLY_JIT_SPLIT=0 keeps
    // 16-byte round trips, BFLY_JIT_SPLIT_MIN_SWEEPS sets the pass size.
    int split_min_sweeps = 6;
    if (const char* ms = std::getenv("BFLY_JIT_SPLIT_MIN_SWEEPS")) split_min_sweeps = std::atoi(ms);
    bool split = !f32 && NT == static_cast<int>(NG) && P.nsweeps > 0 && P.nsweeps >= split_min_sweeps;
    if (const char* sp = std::getenv("BFLY_JIT_SPLIT")) split = split && sp[0] != '0';
    // three waves per SIMD, rounded up to whole blocks (a block of NT threads
    // puts NT / 256 waves on each SIMD)
    const int wave_step = NT >= 256 ? NT / 256 : 1;
    int WPE_used = split ? std::max(WPE, (3 + wave_step - 1) / wave_step * wave_step) : WPE;
    if (const char* w = std::getenv("BFLY_JIT_WAVES_PER_SIMD")) WPE_used = std::max(1, std::atoi(w));
    // Launch form.  One block per tile (split passes; fewer than 4096 tiles;
    // BFLY_JIT_GRID=tile): the hardware dispatcher starts a new block as soon
    // as one finishes.  Work queue (other passes of 4096 tiles or more): a
    // grid of the blocks the device holds at once, each taking its next tile
    // from a counter the executing state holds, with that tile's loads issued
    // during the current tile's first sweep (in the first sweep's register
    // layout where the pass loads directly).  Persistent blocks with a fixed
    // share of the tiles each (BFLY_JIT_GRID=persistent) wait for the slowest
    // block.  Measured on the 30-qubit layered circuit (ms per circuit):
    // 16-byte round trips: complex128 searched / first-fit tile qubits,
    // complex64: queue 282.5-283.1 / 389 / 157.1, one block per tile
    // 285.3-285.5 / 410 / 157.8, fixed shares 320-323, the queue without
    // next-tile loads (BFLY_JIT_PREFETCH=0) 291.4; split complex128: one
    // block per tile 262.0, queue without next-tile loads 273.1, with them
    // 304.9 (their 64 registers spill at three blocks per CU).
    // BFLY_JIT_QUEUE_MIN_TILES sets the threshold (tests).
    unsigned long long min_queue_tiles = 4096;
    if (const char* mt = std::getenv("BFLY_JIT_QUEUE_MIN_TILES")) min_queue_tiles = std::strtoull(mt, nullptr, 10);
    bool queue = !split && (1ull << (P.n - K)) >= min_queue_tiles;
    bool persistent = queue;
    if (const char* g = std::getenv("BFLY_JIT_GRID")) {
        if (g[0] == 't') persistent = queue = false;
        else if (g[0] == 'p') { persistent = true; queue = false; }
        else if (g[0] == 'q') persistent = queue = true;
    }
    // a dispatch covers at most 2^32 - 1 work-items per dimension: beyond
    // that (2^33 amplitudes in 2^9-amplitude tiles, 2^36 in 2^12) the
    // blocks walk the tiles
    if (!persistent &&
        (1ull << (P.n - P.k)) * static_cast<unsigned long long>(tile_threads_h(P.k, P.slots)) > 0xffffffffull)
        persistent = queue = true;
    bool prefetch = persistent && !split;
    if (const char* pf = std::getenv("BFLY_JIT_PREFETCH")) prefetch = persistent && pf[0] == '1';
    const unsigned long long ntiles = 1ull << (P.n - K);

    JitSource js;
    js.threads = NT;
    js.ntiles = ntiles;
    js.persistent = persistent;
    js.queue = queue;
    Out o;
    auto cst = [&](double dv, float fv) {
        const size_t i = f32 ? js.cf.size() : js.cd.size();
        if (f32) js.cf.push_back(fv);
        else js.cd.push_back(dv);
        return static_cast<int>(i);
    };

    int B = 5;
    if (const char* pb = std::getenv("BFLY_JIT_PAD_BITS")) B = std::max(4, std::min(6, std::atoi(pb)));
    const int E = split ? elem : 2 * elem;   // bytes of an LDS element (one component when split)
    // LDS tile: complex-interleaved elements (re, im adjacent: one 16-byte
    // access per complex128 amplitude, 8 bytes per complex64 one), element j
    // at slot pad(j) = j + (j >> B) + (j >> 2B): one spare slot after every
    // 2^B and every 2^2B elements.  Local bit p then moves an element by
    // 2^(p mod B) slots modulo 2^B, so lanes whose index bits have distinct
    // residues mod B fall on distinct banks.  B = 5 for both widths: measured
    // on gfx950 with 64-lane 16-byte accesses, one of ds_read_b128 /
    // ds_write_b128 serves 32 lanes of 8 bytes per cycle over 128 banks and
    // the other 8 lanes of 16 bytes over 32 banks; with B = 5 and thread bits
    // 0-4 on local bits of residues 0-4 both are conflict-free (B = 4: 2.0
    // bank-conflict cycles per access on average in the tile passes of the
    // 30-qubit layered circuit, SQ_LDS_BANK_CONFLICT).  For indices with
    // disjoint bits the map is additive (no carries: pad(a | b) = pad(a) +
    // pad(b)), so a sweep's member addresses are one base register plus
    // immediate offsets; an XOR swizzle needs one XOR per member, which the
    // compiler hoists out of the tile loop for every sweep at once (208
    // VGPRs for 13 sweeps of 16 members: spills).  BFLY_JIT_PAD_BITS=4 keeps
    // the 16-slot rows for measurement; 6 (64-slot rows) measured slower for
    // the one-component round trips too (269 vs 266 ms on the 30-qubit
    // layered circuit).
    auto pad = [&](unsigned j) { return j + (j >> B) + (j >> (2 * B)); };
    const unsigned TSP = pad(TS - 1) + 1;   // slots of the padded tile
    const int VN = f32 ? 4 : 2;   // amplitudes per 16-byte stage-in access

    // thread -> group of a sweep: the first B thread bits go to non-slot local
    // bits of distinct residues mod B (distinct LDS bank groups under pad), the
    // rest to the remaining non-slot bits in ascending order
    auto thread_pos = [&](const TileSweep& sw, int32_t* tpos) {
        bool is_slot[kMaxTileBits] = {};
        for (int a = 0; a < S; ++a) is_slot[sw.s[a]] = true;
        bool used[kMaxTileBits] = {};
        int nt = 0;
        // thread bit i < B goes to the lowest free non-slot bit congruent to
        // i mod B: its slot bit is then i, and the lanes of an LDS pass
        // cover one block of 2^B consecutive slots in order
        for (int i = 0; i < B && i < K - S; ++i) {
            int pick = -1;
            for (int pbit = i; pbit < K && pick < 0; pbit += B)
                if (!is_slot[pbit] && !used[pbit]) pick = pbit;
            if (pick < 0)
                for (int pbit = 0; pbit < K && pick < 0; ++pbit)
                    if (!is_slot[pbit] && !used[pbit]) pick = pbit;
            used[pick] = true;
            tpos[nt++] = pick;
        }
        for (int pbit = 0; pbit < K; ++pbit)
            if (!is_slot[pbit] && !used[pbit]) tpos[nt++] = pbit;
    };
    // Contiguous amplitudes (log2) that one wave-wide access covers: the local
    // bits its 64 lanes and its vector elements span, counted up from local
    // bit 0 while local bit b is global bit b.
    auto contiguous_bits = [&](const bool* covered) {
        int r = 0;
        while (r < K && covered[r] && P.qpos[r] == r) ++r;
        return r;
    };
    auto staged_bits = [&]() {
        bool cov[kMaxTileBits] = {};
        for (int b = 0; b < K && b < 6 + (f32 ? 2 : 1); ++b) cov[b] = true;
        return contiguous_bits(cov);
    };
    auto sweep_bits = [&](const TileSweep& sw) {
        int32_t tpos[kMaxTileBits];
        thread_pos(sw, tpos);
        bool cov[kMaxTileBits] = {};
        for (int i = 0; i < 6 && i < K - S; ++i) cov[tpos[i]] = true;
        bool is_slot[kMaxTileBits] = {};
        for (int a = 0; a < S; ++a) is_slot[sw.s[a]] = true;
        if (is_slot[0]) {
            cov[0] = true;
            if (f32 && is_slot[1] && P.qpos[0] == 0) cov[1] = true;
        }
        return contiguous_bits(cov);
    };
    // Direct stage-in / stage-out: the first sweep loads its members from
    // HBM in its own register layout and the last sweep stores them back, so
    // a pass of s sweeps makes s - 1 LDS round trips instead of s + 1 (a
    // staged pass writes the loaded tile to LDS and reads it back for its
    // first sweep, and the same again before its stores).  A sweep's layout
    // is taken when its wave-wide accesses cover runs of HBM at least as long
    // as the staged layout's (1 KiB when local bits 0-6 are global bits 0-6),
    // and only in passes of 6 or more sweeps: the saved round trips pay in
    // passes bound by their sweeps, while in light, HBM-bound passes the
    // staged layout's access order streams slightly better.  Measured on the
    // 30-qubit layered circuit (ms per circuit, searched / first-fit tile
    // qubits): every pass staged 302 / 416, direct where the runs allow 295 /
    // 422, and in passes of >= 6 sweeps 294.5 / 414.5; every first and last
    // sweep direct whatever its runs 309 / 456.  Needs one group of members
    // per thread.  Measurement switches: BFLY_JIT_DIRECT=0 stages every pass,
    // =1 takes the first and last sweep's layouts whatever their runs, =in /
    // =out only one side; BFLY_JIT_DIRECT_MIN_SWEEPS sets the pass size (1:
    // every pass).
    // Direct member accesses: member m of a thread's group is local index
    // jb | d(m) (disjoint bits), at global offset base + deposit(jb) +
    // deposit(d(m)).  Members differing only in the slots on local bits 0
    // (and 1) sit at consecutive global addresses when those local bits are
    // global bits 0 (and 1): one vector access carries them.
    struct MemberIo {
        int vw;
        int vslot[2];
        unsigned mloc[1 << kSlots];
    };
    auto member_io = [&](const TileSweep& sw) {
        MemberIo f{1, {-1, -1}, {}};
        for (int m = 0; m < G; ++m) {
            unsigned d = 0;
            for (int a = 0; a < S; ++a) d |= ((static_cast<unsigned>(m) >> a) & 1u) << sw.s[a];
            f.mloc[m] = d;
        }
        for (int a = 0; a < S; ++a) {
            if (sw.s[a] == 0 && P.qpos[0] == 0) f.vslot[0] = a;
            if (sw.s[a] == 1 && P.qpos[1] == 1) f.vslot[1] = a;
        }
        if (f.vslot[0] >= 0) f.vw = (f32 && f.vslot[1] >= 0) ? 4 : 2;
        return f;
    };
    // loads (store = false) or stores of a thread's members dr[] / di[] at basevar + gjvar
    auto emit_member_io = [&](const MemberIo& f, bool store, const char* dr, const char* di, const char* basevar,
                              const char* gjvar) {
        static const char* const kComp[4] = {"x", "y", "z", "w"};
        const char* VT = f.vw == 1 ? T : (f32 ? (f.vw == 4 ? "float4" : "float2") : "double2");
        auto vmember = [&](int m, int e) {   // member of element e of the vector at member m
            int r = m;
            if (e & 1) r |= 1 << f.vslot[0];
            if (e & 2) r |= 1 << f.vslot[1];
            return r;
        };
        for (int m = 0; m < G; ++m) {
            if ((f.vw >= 2 && ((m >> f.vslot[0]) & 1)) || (f.vw == 4 && ((m >> f.vslot[1]) & 1))) continue;
            const unsigned long long u = deposit_const(f.mloc[m], P.qpos, K);
            if (f.vw == 1) {
                if (store)
                    o("    re[%s + %s + %lluull] = %s[%d]; im[%s + %s + %lluull] = %s[%d];\n", basevar, gjvar, u, dr, m,
                      basevar, gjvar, u, di, m);
                else
                    o("    %s[%d] = re[%s + %s + %lluull]; %s[%d] = im[%s + %s + %lluull];\n", dr, m, basevar, gjvar, u,
                      di, m, basevar, gjvar, u);
                continue;
            }
            if (store) {
                o("    { %s vr, vi;\n", VT);
                for (int e = 0; e < f.vw; ++e)
                    o("      vr.%s = %s[%d]; vi.%s = %s[%d];\n", kComp[e], dr, vmember(m, e), kComp[e], di, vmember(m, e));
                o("      *reinterpret_cast<%s*>(re + %s + %s + %lluull) = vr; *reinterpret_cast<%s*>(im + %s + %s + %lluull) = vi; }\n",
                  VT, basevar, gjvar, u, VT, basevar, gjvar, u);
            } else {
                o("    { const %s vr = *reinterpret_cast<const %s*>(re + %s + %s + %lluull);"
                  " const %s vi = *reinterpret_cast<const %s*>(im + %s + %s + %lluull);\n",
                  VT, VT, basevar, gjvar, u, VT, VT, basevar, gjvar, u);
                for (int e = 0; e < f.vw; ++e)
                    o("      %s[%d] = vr.%s; %s[%d] = vi.%s;\n", dr, vmember(m, e), kComp[e], di, vmember(m, e), kComp[e]);
                o("    }\n");
            }
        }
    };
    bool direct_in = false, direct_out = false;
    int min_sweeps = 6;
    if (const char* ms = std::getenv("BFLY_JIT_DIRECT_MIN_SWEEPS")) min_sweeps = std::atoi(ms);
    if ((!persistent || (queue && prefetch)) && NT == static_cast<int>(NG) && P.nsweeps > 0 && P.nsweeps >= min_sweeps) {
        const TileSweep& s0 = hp.sweeps[static_cast<size_t>(P.first_sweep)];
        const TileSweep& s1 = hp.sweeps[static_cast<size_t>(P.first_sweep + P.nsweeps - 1)];
        const int need = staged_bits();
        direct_in = sweep_bits(s0) >= need;
        direct_out = sweep_bits(s1) >= need;
        if (const char* d = std::getenv("BFLY_JIT_DIRECT")) {
            if (d[0] == '0') direct_in = direct_out = false;
            else if (d[0] == '1') direct_in = direct_out = true;
            else if (d[0] == 'i') direct_out = false;
            else if (d[0] == 'o') direct_in = false;
        }
    }
    o("typedef %s T;\ntypedef %s C2;\n", T, f32 ? "float2" : "double2");
    o("__device__ __forceinline__ unsigned padc(unsigned j) { return j + (j >> %d) + (j >> %d); }\n", B, 2 * B);
    o("extern \"C\" __global__ __launch_bounds__(%d, %d) void bfly_tile_pass(T* __restrict__ re, T* __restrict__ im, "
      "const T* __restrict__ Cp%s) {\n",
      NT, WPE_used, queue ? ", unsigned long long* __restrict__ Q" : "");
    o("  using namespace bfly;\n");
    if (split) o("  __shared__ __attribute__((aligned(16))) T s_c[%u];\n", TSP);
    else o("  __shared__ __attribute__((aligned(16))) C2 s_c[%u];\n", TSP);
    o("  const unsigned tid = threadIdx.x;\n");
    if (queue) {
        // Work queue: thread 0 takes the next tile index from the pass's
        // counter; the index reaches the block through LDS, double-buffered
        // by iteration parity so that a take never overwrites an index some
        // wave has not read yet (one barrier separates them).  Every block
        // stops at the first index >= ntiles; the counter only grows.
        o("  __shared__ unsigned long long s_tile[2];\n");
        o("  if (tid == 0) s_tile[0] = atomicAdd(Q, 1ull);\n  __syncthreads();\n");
        o("  unsigned long long tile = s_tile[0];\n  if (tile >= %lluull) return;\n  unsigned par = 1u;\n", ntiles);
    } else {
        o("  unsigned long long tile = blockIdx.x;\n  if (tile >= %lluull) return;\n", ntiles);
    }
    // Stage-in / stage-out: each lane moves VN consecutive amplitudes per
    // 16-byte HBM access (re and im); thread tid holds local indices
    // (tid + c NT) VN + v, whose global offsets split into deposit(tid VN) and
    // the constant deposit(c NT VN + v) (disjoint bits).
    const unsigned NV = TS / static_cast<unsigned>(VN);
    const unsigned PV = (NV + static_cast<unsigned>(NT) - 1) / static_cast<unsigned>(NT);
    o("typedef %s V;\n", f32 ? "float4" : "double2");
    o("  const unsigned tv = tid * %du;\n", VN);
    o("  const unsigned long long toff = %s;\n", deposit_expr("tv", P.qpos, 0, K).c_str());
    o("  V pr[%u], pi[%u];\n", PV, PV);
    auto tile_base = [&](const char* var) { return deposit_expr(var, P.opos, 0, P.nouter); };
    auto load = [&](const char* basevar) {
        for (unsigned c = 0; c < PV; ++c) {
            const unsigned long long u = deposit_const(static_cast<unsigned long long>(c) * NT * VN, P.qpos, K);
            const bool guard = (c + 1) * static_cast<unsigned>(NT) > NV;
            if (guard) o("  if (tid + %uu < %uu) {\n", c * NT, NV);
            o("  pr[%u] = *reinterpret_cast<const V*>(re + %s + toff + %lluull);\n", c, basevar, u);
            o("  pi[%u] = *reinterpret_cast<const V*>(im + %s + toff + %lluull);\n", c, basevar, u);
            if (guard) o("  }\n");
        }
    };
    // Split round trips: one component of the staged layout (from the pr or
    // pi registers) into the LDS tile, and a sweep's own layout
    // (thread part, member offsets in bytes) for writing its im parts later.
    auto emit_stage_write = [&](const char* regs) {
        for (unsigned c = 0; c < PV; ++c) {
            if ((c + 1) * static_cast<unsigned>(NT) > NV) o("  if (tid + %uu < %uu)\n", c * NT, NV);
            o("  { const T* a = reinterpret_cast<const T*>(&%s[%u]);\n", regs, c);
            for (int v = 0; v < VN; ++v)
                o("    *reinterpret_cast<T*>(reinterpret_cast<char*>(s_c) + (stv + %uu)) = a[%d];\n",
                  pad(c * NT * VN + static_cast<unsigned>(v)) * E, v);
            o("  }\n");
        }
    };
    auto sweep_layout = [&](int sidx, std::string& jbexpr, unsigned* moff_out) {
        const TileSweep& w = hp.sweeps[static_cast<size_t>(P.first_sweep + sidx)];
        int32_t tp[kMaxTileBits];
        thread_pos(w, tp);
        jbexpr = deposit_expr("tid", tp, 0, K - S);
        for (int m = 0; m < G; ++m) {
            unsigned d = 0;
            for (int a = 0; a < S; ++a) d |= ((static_cast<unsigned>(m) >> a) & 1u) << w.s[a];
            moff_out[m] = pad(d) * E;
        }
    };
    // staged passes with next-tile loads: when split, the staging registers
    // still hold the first tile's im parts until the first sweep has read
    // them back, so the next tile's loads are issued there
    const bool deferred_prefetch = split && persistent && prefetch && !direct_in;
    o("  unsigned long long base = %s;\n", tile_base("tile").c_str());
    // the slot of tv + c NT VN + v (disjoint bits) is pad(tv) + pad(c NT VN + v)
    o("  const unsigned stv = padc(tv) * %du;\n", E);
    if (!direct_in) {
        load("base");
        if (persistent) o("  for (;;) {\n");
        if (queue) o("  if (tid == 0) s_tile[par] = atomicAdd(Q, 1ull);\n");
        if (split) emit_stage_write("pr");
        else
            for (unsigned c = 0; c < PV; ++c) {
                if ((c + 1) * static_cast<unsigned>(NT) > NV) o("  if (tid + %uu < %uu)\n", c * NT, NV);
                o("  { const T* a = reinterpret_cast<const T*>(&pr[%u]); const T* b = reinterpret_cast<const T*>(&pi[%u]);\n", c, c);
                for (int v = 0; v < VN; ++v)
                    o("    { C2 e; e.x = a[%d]; e.y = b[%d]; *reinterpret_cast<C2*>(reinterpret_cast<char*>(s_c) + (stv + %uu)) = e; }\n",
                      v, v, pad(c * NT * VN + static_cast<unsigned>(v)) * E);
                o("  }\n");
            }
        o("  __syncthreads();\n");
    } else if (queue) {
        // the first tile's members in the first sweep's layout; the loop then
        // loads each next tile's during the current tile's first sweep
        const TileSweep& s0 = hp.sweeps[static_cast<size_t>(P.first_sweep)];
        int32_t tp0[kMaxTileBits];
        thread_pos(s0, tp0);
        o("  T nr[%d], ni[%d];\n", G, G);
        o("  {\n    const unsigned jb = static_cast<unsigned>(%s);\n", deposit_expr("tid", tp0, 0, K - S).c_str());
        o("    const unsigned long long gj = %s;\n", deposit_expr("jb", P.qpos, 0, K).c_str());
        emit_member_io(member_io(s0), false, "nr", "ni", "base", "gj");
        o("  }\n  for (;;) {\n");
        o("  if (tid == 0) s_tile[par] = atomicAdd(Q, 1ull);\n  __syncthreads();\n");
    }
    if (!persistent) {
        // one tile per block: nothing to load ahead
    } else if (prefetch && queue && direct_in) {
        o("  const unsigned long long next = s_tile[par];\n  par ^= 1u;\n");
        o("  const unsigned long long nc = next < %lluull ? next : tile;\n", ntiles);
        o("  const unsigned long long nbase = %s;\n", tile_base("nc").c_str());
    } else if (prefetch) {
        if (queue) o("  const unsigned long long next = s_tile[par];\n  par ^= 1u;\n");
        else o("  const unsigned long long next = tile + gridDim.x;\n");
        // The next tile's loads are issued before the sweeps and stay in
        // flight across them.  They are unconditional (the last tile reloads
        // its own amplitudes, which are never used): a conditional load makes
        // the compiler copy the loaded registers where the branches join,
        // which waits for the loads before the first sweep.
        o("  const unsigned long long nc = next < %lluull ? next : tile;\n", ntiles);
        o("  const unsigned long long nbase = %s;\n", tile_base("nc").c_str());
        if (!deferred_prefetch) load("nbase");
    } else {
        if (queue) o("  const unsigned long long next = s_tile[par];\n  par ^= 1u;\n");
        else o("  const unsigned long long next = tile + gridDim.x;\n");
        o("  unsigned long long nbase = 0;\n");
        o("  if (next < %lluull) nbase = %s;\n", ntiles, tile_base("next").c_str());
    }

    // sweeps
    if (split) o("  T ar[%d], ai[%d];\n", G, G);
    for (int s = 0; s < P.nsweeps; ++s) {
        const TileSweep& sw = hp.sweeps[static_cast<size_t>(P.first_sweep + s)];
        o("  // sweep %d: slots %d %d %d %d\n", s, sw.s[0], sw.s[1], S > 2 ? sw.s[2] : -1, S > 3 ? sw.s[3] : -1);
        if (NT > static_cast<int>(NG)) o("  if (tid < %uu)\n", NG);
        int32_t tpos[kMaxTileBits];
        thread_pos(sw, tpos);
        // the sweep's parameters load inside the sweep: an opaque pointer per
        // sweep keeps the compiler from hoisting every constant of the pass out
        // of the tile loop (which spills scalar registers into vector lanes)
        o("  {\n    unsigned zs = 0u;\n    asm volatile(\"\" : \"+s\"(zs));\n    const T* __restrict__ C = Cp + zs;\n");
        o("    const unsigned jb = static_cast<unsigned>(%s);\n", deposit_expr("tid", tpos, 0, K - S).c_str());
        o("    const unsigned jbs = padc(jb) * %du;\n", E);
        if (!split) o("    T ar[%d], ai[%d];\n", G, G);
        unsigned moff[1 << kSlots];
        for (int m = 0; m < G; ++m) {
            unsigned d = 0;
            for (int a = 0; a < S; ++a) d |= ((static_cast<unsigned>(m) >> a) & 1u) << sw.s[a];
            moff[m] = pad(d) * E;
        }
        const bool from_hbm = direct_in && s == 0, to_hbm = direct_out && s == P.nsweeps - 1;
        const MemberIo fio = member_io(sw);
        if (from_hbm || to_hbm) o("    const unsigned long long gj = %s;\n", deposit_expr("jb", P.qpos, 0, K).c_str());
        if (from_hbm && queue) {
            // this tile's members arrived in nr / ni one tile ahead; the next
            // tile's loads go there now and stay in flight across the sweeps
            for (int m = 0; m < G; ++m) o("    ar[%d] = nr[%d]; ai[%d] = ni[%d];\n", m, m, m, m);
            emit_member_io(fio, false, "nr", "ni", "nbase", "gj");
        } else if (from_hbm) {
            emit_member_io(fio, false, "ar", "ai", "base", "gj");
        } else if (split) {
            // re of this layout (written at the previous boundary), then the
            // previous layout's im parts through the same buffer
            for (int m = 0; m < G; ++m)
                o("    ar[%d] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(s_c) + (jbs + %uu));\n", m, moff[m]);
            o("    __syncthreads();\n");
            if (s == 0) {
                emit_stage_write("pi");
            } else {
                std::string jp;
                unsigned mp[1 << kSlots];
                sweep_layout(s - 1, jp, mp);
                o("    { const unsigned jbp = padc(static_cast<unsigned>(%s)) * %du;\n", jp.c_str(), E);
                for (int m = 0; m < G; ++m)
                    o("      *reinterpret_cast<T*>(reinterpret_cast<char*>(s_c) + (jbp + %uu)) = ai[%d];\n", mp[m], m);
                o("    }\n");
            }
            o("    __syncthreads();\n");
            for (int m = 0; m < G; ++m)
                o("    ai[%d] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(s_c) + (jbs + %uu));\n", m, moff[m]);
            if (s == 0 && deferred_prefetch) load("nbase");
        } else {
            for (int m = 0; m < G; ++m)
                o("    { const C2 e = *reinterpret_cast<const C2*>(reinterpret_cast<const char*>(s_c) + (jbs + %uu)); ar[%d] = e.x; ai[%d] = e.y; }\n",
                  moff[m], m, m);
        }
        std::vector<IrOp> ir;
        for (int q = 0; q < sw.ninstr; ++q)
            lower_instr(ir, hp.instrs[static_cast<size_t>(sw.first + q)], S, f32, cst);
        fold_signs(ir, S);
        emit_ops(o, ir, S, cst);
        if (to_hbm) {
            emit_member_io(fio, true, "ar", "ai", "base", "gj");
            o("  }\n");
        } else if (split) {
            for (int m = 0; m < G; ++m)
                o("    *reinterpret_cast<T*>(reinterpret_cast<char*>(s_c) + (jbs + %uu)) = ar[%d];\n", moff[m], m);
            o("  }\n  __syncthreads();\n");
        } else {
            for (int m = 0; m < G; ++m)
                o("    { C2 e; e.x = ar[%d]; e.y = ai[%d]; *reinterpret_cast<C2*>(reinterpret_cast<char*>(s_c) + (jbs + %uu)) = e; }\n",
                  m, m, moff[m]);
            o("  }\n  __syncthreads();\n");
        }
    }

    // stage out
    if (!direct_out && split) {
        // re from the staged layout, then the last sweep's im parts through
        // the same buffer
        o("  { V sr[%u];\n", PV);
        for (unsigned c = 0; c < PV; ++c) {
            if ((c + 1) * static_cast<unsigned>(NT) > NV) o("  if (tid + %uu < %uu)\n", c * NT, NV);
            o("  { T* a = reinterpret_cast<T*>(&sr[%u]);\n", c);
            for (int v = 0; v < VN; ++v)
                o("    a[%d] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(s_c) + (stv + %uu));\n", v,
                  pad(c * NT * VN + static_cast<unsigned>(v)) * E);
            o("  }\n");
        }
        o("  __syncthreads();\n");
        {
            std::string jp;
            unsigned mp[1 << kSlots];
            sweep_layout(P.nsweeps - 1, jp, mp);
            o("  { const unsigned jbp = padc(static_cast<unsigned>(%s)) * %du;\n", jp.c_str(), E);
            for (int m = 0; m < G; ++m)
                o("    *reinterpret_cast<T*>(reinterpret_cast<char*>(s_c) + (jbp + %uu)) = ai[%d];\n", mp[m], m);
            o("  }\n");
        }
        o("  __syncthreads();\n");
        for (unsigned c = 0; c < PV; ++c) {
            const unsigned long long u = deposit_const(static_cast<unsigned long long>(c) * NT * VN, P.qpos, K);
            if ((c + 1) * static_cast<unsigned>(NT) > NV) o("  if (tid + %uu < %uu)\n", c * NT, NV);
            o("  { V vi; T* b = reinterpret_cast<T*>(&vi);\n");
            for (int v = 0; v < VN; ++v)
                o("    b[%d] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(s_c) + (stv + %uu));\n", v,
                  pad(c * NT * VN + static_cast<unsigned>(v)) * E);
            o("    *reinterpret_cast<V*>(re + base + toff + %lluull) = sr[%u]; *reinterpret_cast<V*>(im + base + toff + %lluull) = vi; }\n",
              u, c, u);
        }
        o("  }\n");
    } else if (!direct_out) {
        for (unsigned c = 0; c < PV; ++c) {
            const unsigned long long u = deposit_const(static_cast<unsigned long long>(c) * NT * VN, P.qpos, K);
            if ((c + 1) * static_cast<unsigned>(NT) > NV) o("  if (tid + %uu < %uu)\n", c * NT, NV);
            o("  { V vr, vi; T* a = reinterpret_cast<T*>(&vr); T* b = reinterpret_cast<T*>(&vi);\n");
            for (int v = 0; v < VN; ++v)
                o("    { const C2 e = *reinterpret_cast<const C2*>(reinterpret_cast<const char*>(s_c) + (stv + %uu)); a[%d] = e.x; b[%d] = e.y; }\n",
                  pad(c * NT * VN + static_cast<unsigned>(v)) * E, v, v);
            o("    *reinterpret_cast<V*>(re + base + toff + %lluull) = vr; *reinterpret_cast<V*>(im + base + toff + %lluull) = vi; }\n",
              u, u);
        }
    }
    if (persistent) {
        o("  if (next >= %lluull) break;\n  __syncthreads();\n  tile = next;\n  base = nbase;\n", ntiles);
        if (!prefetch) load("base");
        o("  }\n");
    }
    o("}\n");
    js.text = std::move(o.s);
    return js;
}

namespace {

std::mutex g_mu;
std::unordered_map<std::string, std::shared_ptr<JitCode>> g_codes;       // by source text
std::map<std::pair<const JitCode*, int>, hipFunction_t> g_functions;      // by (code, device)

const char* offload_arch() {
    const char* a = std::getenv("BFLY_OFFLOAD_ARCH");
    return a && *a ? a : "gfx950";
}

// Compile one source; returns an empty pointer and fills err on failure.
std::shared_ptr<JitCode> compile_source(const std::string& text, std::string& err) {
    hiprtcProgram prog = nullptr;
    const char* hdr_src[] = {kTileOps};
    const char* hdr_names[] = {"tile_ops.h"};
    const std::string full = std::string("#include \"tile_ops.h\"\n") + text;
    hiprtcResult r = hiprtcCreateProgram(&prog, full.c_str(), "bfly_tile_pass.hip", 1, hdr_src, hdr_names);
    if (r != HIPRTC_SUCCESS) {
        err = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
        return nullptr;
    }
    const std::string arch = std::string("--offload-arch=") + offload_arch();
    // packed FP32 instructions stay disabled, as for the library's own kernels
    const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    r = hiprtcCompileProgram(prog, 7, opts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        err = std::string("hiprtcCompileProgram: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 4000);
        hiprtcDestroyProgram(&prog);
        return nullptr;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    auto co = std::make_shared<JitCode>();
    co->code.resize(cs);
    hiprtcGetCode(prog, co->code.data());
    hiprtcDestroyProgram(&prog);
    return co;
}

}  // namespace

int jit_build(JitPlan& jp, const HostPlan& hp, std::string& err) {
    jp = JitPlan{};
    jp.step_pass.assign(hp.steps.size(), -1);
    std::vector<JitSource> srcs;
    for (size_t i = 0; i < hp.steps.size(); ++i) {
        if (!jit_supported(hp, hp.steps[i])) continue;
        jp.step_pass[i] = static_cast<int>(jp.passes.size());
        srcs.push_back(jit_generate(hp, hp.steps[i]));
        JitPass pass;
        pass.threads = srcs.back().threads;
        pass.ntiles = srcs.back().ntiles;
        pass.persistent = srcs.back().persistent;
        pass.queue = srcs.back().queue;
        if (pass.queue) pass.qidx = jp.nqueue++;
        // constants: 16-byte aligned offsets into one array per plan
        const size_t elem = hp.dtype == BFLY_C64 ? 4 : 8;
        while ((jp.consts.size() % 16) != 0) jp.consts.push_back(0);
        pass.const_off = jp.consts.size();
        const char* b = hp.dtype == BFLY_C64 ? reinterpret_cast<const char*>(srcs.back().cf.data())
                                             : reinterpret_cast<const char*>(srcs.back().cd.data());
        const size_t nb = (hp.dtype == BFLY_C64 ? srcs.back().cf.size() : srcs.back().cd.size()) * elem;
        jp.consts.insert(jp.consts.end(), b, b + nb);
        jp.passes.push_back(pass);
    }
    // BFLY_JIT_DUMP=<dir>: write every generated source there (inspection)
    if (const char* dump = std::getenv("BFLY_JIT_DUMP")) {
        for (size_t i = 0; i < srcs.size(); ++i) {
            const std::string path = std::string(dump) + "/pass" + std::to_string(i) + ".hip";
            if (FILE* f = std::fopen(path.c_str(), "w")) {
                std::fputs("#include \"tile_ops.h\"\n", f);
                std::fputs(srcs[i].text.c_str(), f);
                std::fclose(f);
            }
        }
    }
    // compile what the cache lacks, in parallel
    std::vector<size_t> todo;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (size_t i = 0; i < srcs.size(); ++i) {
            auto it = g_codes.find(srcs[i].text);
            if (it != g_codes.end()) jp.passes[i].code = it->second;
            else todo.push_back(i);
        }
    }
    jp.compiled = static_cast<int>(todo.size());
    if (!todo.empty()) {
        std::atomic<size_t> next{0};
        std::vector<std::string> errs(todo.size());
        unsigned hw = std::thread::hardware_concurrency();
        const unsigned nthreads = std::max(1u, std::min<unsigned>({hw ? hw : 4u, 16u, static_cast<unsigned>(todo.size())}));
        auto work = [&]() {
            for (;;) {
                const size_t w = next.fetch_add(1);
                if (w >= todo.size()) return;
                const size_t i = todo[w];
                jp.passes[i].code = compile_source(srcs[i].text, errs[w]);
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nthreads; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
        for (size_t w = 0; w < todo.size(); ++w) {
            if (!jp.passes[todo[w]].code) {
                err = "tile-pass kernel compile failed: " + errs[w];
                return BFLY_EUNSUPPORTED;
            }
        }
        std::lock_guard<std::mutex> lk(g_mu);
        for (size_t w = 0; w < todo.size(); ++w) {
            const size_t i = todo[w];
            auto ins = g_codes.emplace(srcs[i].text, jp.passes[i].code);
            jp.passes[i].code = ins.first->second;   // share the cached object if another thread won
        }
    }
    jp.built = true;
    return BFLY_OK;
}

int jit_load(JitPlan& jp, int device, std::string& err) {
    if (jp.device == device) return BFLY_OK;
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        err = std::string("hipDeviceGetAttribute: ") + hipGetErrorString(e);
        return BFLY_EHIP;
    }
    for (JitPass& ps : jp.passes) {
        hipFunction_t fn = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            auto key = std::make_pair(static_cast<const JitCode*>(ps.code.get()), device);
            auto it = g_functions.find(key);
            if (it != g_functions.end()) {
                fn = it->second;
            } else {
                // modules stay loaded until the process ends (the cache owns them)
                hipModule_t mod = nullptr;
                e = hipModuleLoadData(&mod, ps.code->code.data());
                if (e == hipSuccess) e = hipModuleGetFunction(&fn, mod, "bfly_tile_pass");
                if (e != hipSuccess) {
                    err = std::string("loading a tile-pass kernel: ") + hipGetErrorString(e);
                    return BFLY_EHIP;
                }
                g_functions.emplace(key, fn);
            }
        }
        int per_cu = 0;
        e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, ps.threads, 0);
        if (e != hipSuccess) {
            err = std::string("hipModuleOccupancyMaxActiveBlocksPerMultiprocessor: ") + hipGetErrorString(e);
            return BFLY_EHIP;
        }
        const unsigned long long cap = static_cast<unsigned long long>(per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
        // one block per tile (fewer than 2^32 work-items: the generator
        // builds the persistent form beyond that), or a persistent grid of
        // the blocks the device holds at once
        ps.grid = static_cast<unsigned>(ps.persistent ? (ps.ntiles < cap ? ps.ntiles : cap) : ps.ntiles);
        ps.fn = fn;
    }
    jp.device = device;
    return BFLY_OK;
}

hipError_t jit_launch(const JitPass& ps, void* re, void* im, const void* consts, unsigned long long* counter,
                      hipStream_t stream) {
    void* c = const_cast<void*>(consts);
    void* args[4] = {&re, &im, &c, &counter};
    return hipModuleLaunchKernel(ps.fn, ps.grid, 1, 1, static_cast<unsigned>(ps.threads), 1, 1, 0, stream, args,
                                 nullptr);
}

int jit_cache_size() {
    std::lock_guard<std::mutex> lk(g_mu);
    return static_cast<int>(g_codes.size());
}

}  // namespace bfly
