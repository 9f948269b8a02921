// kernels.h -- launch wrappers of the CDNA4 kernels (kernels.hip), called by
// the C-ABI layer (api.hip).  Every wrapper enqueues on `stream` and returns
// the hipError_t of the launch.
#pragma once
#include <hip/hip_runtime.h>

#include "bfly_plan.h"

namespace bfly {

// (b) one tile pass: every gate of the pass in one HBM read + write.
hipError_t launch_tile_pass(int dtype, void* re, void* im, const TilePass& P,
                            const TileSweep* d_sweeps, const TileInstr* d_instrs, hipStream_t stream);
// One circuit's tile pass inside a batched launch (device-resident array).
struct BatchItem {
    void* re;
    void* im;
    const TileSweep* sweeps;
    const TileInstr* instrs;
    uint64_t tile_begin;   // first global tile of this item in the launch
    TilePass P;
};
// (b') the next tile pass of many circuits in one persistent launch; all
// items share the tile size k, the sweep width and the dtype.  d_items == nullptr only
// prepares the kernel (attributes, occupancy) without launching.
hipError_t launch_tile_batch(int dtype, int k, int slots, const BatchItem* d_items, int nitems, uint64_t total_tiles,
                             hipStream_t stream);
// Sets kernel attributes and sizes the persistent grids of the tile kernels
// for (dtype, k, slots) outside any stream capture.
hipError_t prepare_tile_kernels(int dtype, int k, int slots);
// (b'') BLOCK mode pass (complex64): a 32 x 32 block on FP32 MFMA between
// FP64-generated phase factors; `mat` is the block's 64 x 64 real form,
// `tab` the plan's column factors and outer-bit angles, `norm2` the squared
// norm to divide out (B.renorm) left on the device by a reduction.
hipError_t launch_block_pass(int dtype, void* re, void* im, const TilePass& P, const BlockPass& B, const float* mat,
                             const double* tab, const double* norm2, hipStream_t stream);
// (a) one gate over the whole state (coalesced strided butterfly pairs).
hipError_t launch_gate(int dtype, void* re, void* im, const GateLaunch& G, hipStream_t stream);
// (c) bit reversal of all bits (mask 0) or of the bits in `mask`, in place.
hipError_t launch_bitrev(int dtype, void* re, void* im, int n, uint64_t mask, hipStream_t stream);
hipError_t launch_swap(int dtype, void* re, void* im, int n, int a, int b, hipStream_t stream);
// dense 2^m block on [target, target+m), udata complex interleaved FP64.
hipError_t launch_unitary(int dtype, void* re, void* im, int n, int target, int m, uint64_t controls,
                          const double* d_u, hipStream_t stream);
hipError_t launch_reset(int dtype, void* re, void* im, uint64_t size, hipStream_t stream);
hipError_t launch_zero_words(unsigned long long* p, int n, hipStream_t stream);
hipError_t launch_scale(int dtype, void* re, void* im, uint64_t size, double scale, hipStream_t stream);
// Measurement (apply_measure, state.mojo:447-553): outcome probabilities of
// up to kOutcomeMaxBits free qubits among the amplitudes with the forced
// values, FP64 compensated and deterministic; `part` holds nblocks * 2^f
// doubles (outcome_blocks), `out` 2^f.
constexpr int kOutcomeMaxBits = 12;
struct OutcomeSpec {
    int32_t f;                 // free qubits (outcome bit k = qubit fpos[k])
    int32_t pad;
    int32_t fpos[kOutcomeMaxBits];
    uint64_t fval;             // values of the forced qubits (their bits set where forced to 1)
    uint64_t rmask;            // every other qubit
    uint64_t rest_per_block;   // rest values per block
};
int outcome_blocks(const OutcomeSpec& sp, uint64_t rest_count);
// More than kOutcomeMaxBits (and at most kOutcomeDirectMaxBits) distinct free
// qubits: one thread per outcome and rest chunk sums its amplitudes (rest
// values in increasing order, TwoSum in FP64), then the chunks are added in
// order: deterministic.
constexpr int kOutcomeDirectMaxBits = 24;
struct OutcomeDirect {
    int32_t f;                 // free qubits (outcome bit k = qubit fpos[k])
    int32_t pad;
    int32_t fpos[kOutcomeDirectMaxBits];
    uint64_t fval;             // forced bits
    uint64_t rmask;            // every other qubit
    uint64_t nrest;            // 2^popcount(rmask)
};
int outcome_direct_chunks(const OutcomeDirect& sp);   // R, a power of two dividing nrest
// part: R * 2^f doubles (unused when R == 1); out: 2^f doubles
hipError_t launch_outcome_direct(int dtype, const void* re, const void* im, const OutcomeDirect& sp, int R,
                                 double* part, double* out, hipStream_t stream);
hipError_t launch_outcome_probs(int dtype, const void* re, const void* im, const OutcomeSpec& sp, int nblocks,
                                double* part, double* out, hipStream_t stream);
// Amplitudes with (i & tmask) == tval times inv (in the state's precision), the others 0.
hipError_t launch_collapse(int dtype, void* re, void* im, uint64_t size, uint64_t tmask, uint64_t tval, double inv,
                           hipStream_t stream);
// Element-wise precision conversion from a src_dtype state to the other dtype.
hipError_t launch_convert(int src_dtype, const void* sre, const void* sim, void* dre, void* dim, uint64_t size,
                          hipStream_t stream);
// Scales to unit norm using the squared norm at d_norm2 (on the device).
hipError_t launch_normalize(int dtype, void* re, void* im, uint64_t size, const double* d_norm2, hipStream_t stream);
// re, im ~ N(0,1) from a SplitMix64 stream (not normalised).
hipError_t launch_random_fill(int dtype, void* re, void* im, uint64_t size, uint64_t seed, hipStream_t stream);

// One state's share of a phase-flip or whole-state diffuser launch (alone, or
// batched over many states).
struct OpItem {
    void* re;
    void* im;
    uint64_t others;      // flip: bits outside the mask
    uint64_t value;       // flip: pattern
    uint64_t count;       // flip: amplitudes negated; diffuse: chunks
    uint64_t begin;       // first work unit of this item in the launch
    double* partial;      // diffuse: 4 doubles per chunk (the state's scratch)
    int32_t chunk_bits;   // diffuse: log2 of the amplitudes per chunk
    int32_t variant;      // diffuse: 0 gate action, 1 the reference's classical update
    double* partial_next; // diffuse chains: where the next diffuser's partials go
    uint64_t fmask;       // diffuse chains: the phase flip applied after the update
    uint64_t fvalue;
    int32_t flags;        // kChain* below
    int32_t pad;
};
OpItem flip_item(void* re, void* im, int n, uint64_t mask, uint64_t value);
OpItem diffuse_item(void* re, void* im, int n, int variant, double* scratch, int32_t chain = 0,
                    uint64_t fmask = 0, uint64_t fvalue = 0);
// PHASE_FLIP: negate every amplitude with (i & mask) == value.  d_items ==
// nullptr runs `one` alone.
hipError_t launch_flip_items(int dtype, const OpItem& one, const OpItem* d_items, int nitems, uint64_t total,
                             hipStream_t stream);
// DIFFUSE over every qubit: compensated chunk sums, then the update of every
// chunk from the same ordered total (identical bits alone or batched).
hipError_t launch_diffuse_items(int dtype, const OpItem& one, const OpItem* d_items, int nitems, uint64_t total,
                                hipStream_t stream);
hipError_t launch_flip(int dtype, void* re, void* im, int n, uint64_t mask, uint64_t value, hipStream_t stream);
// DIFFUSE over the bits of `mask` (every qubit, or at most kDiffuseMaxBits);
// uses `scratch` (reduce_scratch_doubles()) when the mask covers every qubit.
hipError_t launch_diffuse(int dtype, void* re, void* im, int n, uint64_t mask, int variant, double* scratch,
                          int32_t chain, uint64_t fmask, uint64_t fvalue,
                          hipStream_t stream);

// (d) reductions.  Partials live in `scratch` (at least reduce_scratch_doubles()
// doubles); the final value(s) are left in scratch[0..nout) on the device.
enum ReduceMode : int { R_NORM2 = 0, R_PROB_MASK = 1, R_DIFF2 = 2, R_INNER = 3 };
size_t reduce_scratch_doubles();
hipError_t launch_reduce(int dtype, int mode, const void* are, const void* aim, const void* bre,
                         const void* bim, uint64_t size, uint64_t mask, uint64_t value,
                         double* scratch, hipStream_t stream);
// The same reduction with the per-block partials in `partial` (2 * channels *
// reduce_blocks(size) doubles) and the result(s) written to `out` (device).
hipError_t launch_reduce_into(int dtype, int mode, const void* are, const void* aim, const void* bre, const void* bim,
                              uint64_t size, uint64_t mask, uint64_t value, double* partial, double* out,
                              hipStream_t stream);
unsigned reduce_blocks(uint64_t size);
// marginal distribution over m <= 6 target bits; 2^m results in scratch[0..)
hipError_t launch_marginal(int dtype, const void* re, const void* im, uint64_t size, const int* targets,
                           int m, double* scratch, hipStream_t stream);

}  // namespace bfly
