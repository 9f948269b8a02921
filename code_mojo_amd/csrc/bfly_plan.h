// bfly_plan.h -- execution-plan records shared by the host planner
// (planner.cpp) and the HIP kernels (kernels.hip).  Plain C++ PODs: no HIP
// types, so the planner compiles with the host compiler alone.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bfly_hip.h"

namespace bfly {

// Largest tile (2^k amplitudes staged in LDS) and the widest state handled.
constexpr int kMaxTileBits = 13;
constexpr int kMinTileBits = 4;
constexpr int kMaxQubits = 40;
// Widest diffuser block handled by one workgroup (wider ones must cover every qubit).
constexpr int kDiffuseMaxBits = 12;

// Gate application kinds inside a tile sweep or a stand-alone gate launch.
// The specialised kinds reproduce the reference's specialised scalar kernels
// (transformations_scalar.mojo:178-394,397-566); GEN is process_pair_scalar.
enum InstrKind : int32_t {
    I_GEN = 0,     // generic 2x2 on a pair
    I_H = 1,       // (u +- v) * 0.7071067811865476
    I_X = 2,       // exchange
    I_RY = 3,      // real rotation: m[0] = cos(t/2), m[1] = sin(t/2)
    I_DIAG = 4,    // a *= d0 (target bit 0) or d1 (target bit 1): m[0..1] = d0, m[2..3] = d1
    I_PHASE1 = 5,  // a *= d1 where target bit is 1 only (P, controlled P)
    I_NEG1 = 6,    // a = -a where target bit is 1 (Z, CZ, multi-controlled Z): exact
};

// Instruction kinds inside a tile sweep (register form of the gates above):
// every pair gate is a 2x2 matrix on one slot; diagonal gates either flip
// signs (Z family, exact) or multiply by a phase.
enum TileKind : int32_t {
    T_PAIR = 0,    // m[0..7]: 2x2 matrix on slot `slot`
    T_NEG = 1,     // negate the members selected by mm1 (target bit 1) or mm0 (target bit 0;
                   // phase flips whose pattern has a 0 at the target)
    T_DIAG = 2,    // multiply members selected by mm1 by d1 = m[2..3], by mm0 by d0 = m[0..1]
    // Factored single-qubit unitaries (no controls): U = sigma * PH(e^{i phi}) *
    // R * PH(e^{i lambda}) with R real; the pass's sigmas are applied by T_SCALE.
    T_ROT = 3,     // on slot `slot`: v <- v * m[2..3] (pre-phase), then the scaled real
                   // rotation with t = m[0], form = m[1]:
                   // form 0: (u, v) <- (u - t v, v + t u); form 1: (t u - v, u + t v)
    T_PH = 4,      // members whose slot bit is 1 times m[0..1]
    T_SCALE = 5,   // every member times m[0..1]
    T_ROTS = 6,    // (3-slot sweeps) T_ROT on every slot s in the mask `cslots`, in slot order (they
                   // commute): m[3s] = t (negative, or -0.0, for form 1),
                   // m[3s + 1 .. 3s + 2] = pre-phase
    T_ROT3 = 7,    // (3-slot sweeps) T_ROTS on all three slots, every rotation in form 0
                   // (straight-line code: no per-slot or per-form branches)
};

// Amplitudes per thread in a tile sweep: 2^S, spanned by S local bits (the
// slots); S = 4 (16 amplitudes) or 3 (8 amplitudes, twice the threads per
// tile, tiles of 2^10 amplitudes or more).
constexpr int kSlots = 4;            // widest sweep
constexpr int kMinSlots3Bits = 10;   // smallest tile with 3-slot sweeps
constexpr int kGroup = 1 << kSlots;  // member masks are kGroup bits wide

// One gate inside a tile pass.  240 bytes; the first 20 bytes and the
// parameters are what every sweep reads (the outer-control words only when
// `guard` says so).
struct alignas(16) TileInstr {
    int8_t kind;       // TileKind
    int8_t slot;       // target's index in the sweep's slot set, or -1 (diagonal
                       // ops only: target outside the slots)
    uint8_t cslots;    // control bits that are slots, as a mask over slot indices
    int8_t dlocal;     // diagonal op with target local but not a slot: its local bit, else -1
    int8_t obitpos;    // diagonal op with target outside the tile: its global bit, else -1
    uint8_t guard;     // 1: outer controls (omask / oval) must be tested
    uint16_t pad0;
    uint16_t mm1;      // diagonal ops: members whose slot controls hold and whose
                       // target bit is 1 (every slot-control member when the target is no slot)
    uint16_t mm0;      // diagonal ops: the same with target bit 0
    uint32_t lmask;    // control bits local to the tile but not slots (local-index bits)
    uint32_t lval;     // required values of the lmask bits (lmask for ordinary controls)
    uint32_t pad1[3];
    uint64_t omask;    // control bits outside the tile, as global-index bits
    uint64_t oval;     // required values of the omask bits (omask for ordinary controls)
    double m[16];      // matrix / parameters in FP64 (T_ROTS uses 12)
    float mf[16];      // the same in FP32 (complex64 states)
};
static_assert(sizeof(TileInstr) == 240, "TileInstr layout");

// A sweep: every group of 16 amplitudes spanned by the kSlots local bits
// `s` is loaded from LDS into registers, all the sweep's gates run in
// registers, and the group is written back.  32 bytes.
// XOR swizzle of a tile's LDS index (elements of `elem_bytes` bytes): bits
// 4.. (8-byte elements) or 5.. (4-byte elements) are folded onto the low bits,
// which spreads the strided accesses of a sweep over the LDS banks.  Linear
// over GF(2): swz(a | b) = swz(a) ^ swz(b) for disjoint a, b.
constexpr uint32_t lds_swizzle(uint32_t j, int elem_bytes) {
    return elem_bytes == 8 ? j ^ ((j >> 4) & 15u) : j ^ ((j >> 5) & 31u);
}

struct alignas(16) TileSweep {
    int32_t s[kSlots]; // local bit positions, distinct (the first S are used)
    int32_t ninstr;
    int32_t first;     // index of the first instruction
    int32_t pad[2];
    // byte offset in an LDS tile of member m's slot bits, swizzled:
    // lds_swizzle(sum over slots a of bit a of m << s[a]) * element size; a
    // member's address is the thread's own swizzled offset XOR doff[m]
    uint32_t doff[kGroup];
};
static_assert(sizeof(TileSweep) == 96, "TileSweep layout");

// Launch descriptor of one tile pass (passed by value as a kernel argument).
struct TilePass {
    int32_t n;
    int32_t k;          // tile bits
    int32_t slots;      // sweep width S (3 or 4)
    int32_t nsweeps;
    int32_t first_sweep;
    int32_t qpos[kMaxTileBits];   // global bit of local bit i (ascending)
    int32_t nouter;
    int32_t opos[kMaxQubits];     // global bits outside the tile (ascending)
};

// A stand-alone single-gate launch (unfused mode, and tiny states).
struct GateLaunch {
    int32_t kind;       // InstrKind
    int32_t target;
    int32_t npos;       // number of inserted-zero positions (target + controls)
    int32_t pos[kMaxQubits + 1];  // ascending positions where zero bits are inserted
    uint64_t cmask;     // control mask (global bits)
    uint64_t count;     // work items: pairs (or single amplitudes for diagonal kinds)
    double m[8];
    float mf[8];
};

enum StepKind : int32_t {
    S_TILE = 0, S_GATE = 1, S_BITREV = 2, S_SWAP = 3, S_UNITARY = 4, S_FLIP = 5, S_DIFFUSE = 6,
    S_BLOCK = 7,   // dense 2^5-amplitude block on FP32 MFMA (complex64 BLOCK mode)
    S_NORM = 8,    // compensated squared norm into the state's scratch (BLOCK mode check)
    S_MEASURE = 9, // measurement: a = targets, uoff = their (qubit, value) list, b = 1 when `controls` is the seed
};

// BLOCK mode: a block is a 32 x 32 complex unitary U on 5 qubits (the
// block's slots), applied to every group of 32 amplitudes that share the
// other bits, between phase factors from controlled-phase gates that couple
// at most one slot to qubits outside the block: amplitude i is multiplied by
// exp(i theta) when (i & omask) == omask and, for slot >= 0, the slot bit is 1.
constexpr int kBlockQubits = 5;
constexpr int kBlockDim = 1 << kBlockQubits;
struct BlockPhase {     // planner record of one phase gate
    uint64_t omask;    // bits outside the block that must be 1 (global index bits; at most one)
    int32_t slot;      // block slot that must be 1, or -1
    int32_t pad;
    double theta;      // radians, FP64 (exact for the dyadic pi / 2^k ladders)
};

// Phase records of a block, pre (applied before U) and post, summed per
// table row in FP64: row 0 is constant, rows 1 .. k - 5 belong to the
// column bits (ascending local position), the rest to the outer bits (tile
// index bits); a row adds its 6 angles (c0 and one per slot) where its bit
// is 1.  Amplitude e of a column then takes exp(i (c0 + sum_j c_j e_j)).
// The device gets, per block, the column factors exp(i c) of rows 0 and the
// column bits as [2][2^(k-5)][6] complex FP64 (FP64 sincos of FP64 sums),
// then the outer-bit angles as [2][n - k][6] FP64.
struct BlockPass {
    int32_t qloc[kBlockQubits];  // tile-local bit of each slot
    int32_t renorm;              // 1: scale by 1 / sqrt(scratch[0]) on load
    uint64_t mat_off;            // first float of the 64 x 64 real form of U in HostPlan::bmat
    uint64_t tab_off;            // first double of the block's device tables in HostPlan::btab
    uint64_t rows_off;           // first double of the block's row angles in HostPlan::brows
    int32_t pre_tab, post_tab;   // 1: that side has a nonzero angle
};

// Grover chains D [F] D: the apply pass of a whole-state diffuser can also
// apply the phase flip that follows it and leave the chunk sums the next
// diffuser needs (the same per-thread order as its own partial pass:
// identical bits), so the next diffuser skips its partial pass.  Partials
// alternate between two halves of the state's scratch.
constexpr int32_t kChainNext = 1;    // apply pass writes the next partials
constexpr int32_t kChainReady = 2;   // partials already written by the previous diffuser
constexpr int32_t kChainOdd = 4;     // this diffuser's partials are in the second half
constexpr int32_t kChainFlip = 8;    // apply pass applies the flip (fmask, fvalue)
constexpr uint64_t kPartialSpan = 4ull << 14;   // doubles per partial buffer (2^14 chunks)

struct Step {
    StepKind kind;
    TilePass tile;       // S_TILE, S_BLOCK (tile bits = the block's slots + column bits)
    BlockPass block;     // S_BLOCK
    GateLaunch gate;     // S_GATE
    uint64_t mask = 0;   // S_BITREV: target mask (0 = all); S_FLIP / S_DIFFUSE: bit mask
    int32_t a = 0, b = 0;  // S_SWAP qubits; S_UNITARY: a = target, b = m; S_DIFFUSE: a = variant
    uint64_t controls = 0; // S_UNITARY control mask; S_FLIP: pattern
    uint64_t uoff = 0;     // S_UNITARY: complex offset into the plan's unitary data
    int32_t gates = 0;     // S_TILE: gates the pass applies (cost model of BFLY_MODE_AUTO)
    int32_t chain = 0;     // S_DIFFUSE over every qubit: kChain* flags (Grover chains)
    uint64_t fmask = 0, fvalue = 0;  // S_DIFFUSE with kChainFlip: the absorbed phase flip
};

struct HostPlan {
    int32_t n = 0;
    int32_t dtype = BFLY_C128;
    bfly_exec_opts opts{};
    std::vector<Step> steps;
    std::vector<TileSweep> sweeps;
    std::vector<TileInstr> instrs;
    std::vector<double> udata;   // complex, interleaved
    std::vector<float> bmat;     // BLOCK mode, 4096 words per block, row-major: complex64 the 64 x 64
                                 // FP32 real form of U, complex128 the 32 x 64 FP64 half [Ur, -Ui]
    std::vector<double> btab;          // BLOCK mode device tables (column factors, outer angles)
    int64_t measures = 0;              // S_MEASURE steps (they synchronise with the host)
    std::vector<double> brows;         // BLOCK mode row angles [2][n - 4][6] per block (host only)
    int64_t gates = 0;
    int64_t passes = 0;          // HBM sweeps over the state
};

// Builds `plan` from validated ops.  Returns BFLY_OK or an error code and
// fills `err`.
int build_plan(HostPlan& plan, int32_t n, int32_t dtype, const bfly_op* ops, uint64_t nops,
               const double* udata, uint64_t udata_len, const bfly_exec_opts& opts, std::string& err);

// validate_circuit (butterfly/core/executors.mojo:129-185) for lowered ops.
int validate_ops(int32_t n, const bfly_op* ops, uint64_t nops, uint64_t udata_len, std::string& err);

// Default tile / low-bit choices for a dtype (also used by the autotuner).
int default_tile_bits(int32_t dtype);
int default_low_bits(int32_t dtype);
int default_sweep_slots(int32_t dtype, int k);

}  // namespace bfly
