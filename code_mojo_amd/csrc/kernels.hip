// kernels.hip -- hand-written CDNA4 (gfx950) kernels of the statevector engine.
//
// State layout in HBM: SoA, re[2^n] and im[2^n] of T (double for complex128,
// float for complex64), exactly the reference's State (state.mojo:19-33), so
// host copies are plain memcpys.
//
// Kernels:
//   (a) gate_kernel       one 2x2 gate, pairs enumerated by inserting zero bits
//                         at the target and control positions (only matching
//                         pairs are read: a multi-controlled gate touches
//                         2^(n-1-c) pairs, not the reference's 2^(n-1) scan,
//                         transformations_scalar.mojo:586-594)
//   (b) tile_pass_kernel  a 2^K-amplitude tile staged in LDS with coalesced
//                         16-byte loads; the pass's gates run as register
//                         sweeps over 8- or 16-amplitude groups; one HBM read + write
//                         per pass (generalises execute_grid_fused's row-local
//                         groups, executors.mojo:1062-1093)
//   (c) bitrev kernels    in place: 32x32 LDS tiles pair index (h,m,l) with
//                         (rev l, rev m, rev h), re and im together, both sides
//                         coalesced (bit_reverse_state, state.mojo:131-204, is
//                         out of place)
//   (d) reduce kernels    FP64 TwoSum-compensated sums with a fixed reduction
//                         tree (deterministic), for norms / probabilities /
//                         inner products (grover_classical.mojo:8-31,109-131)
//   block_mfma_kernel     complex64 dense 5-qubit blocks on FP32 MFMA with FP64
//                         phase factors (BLOCK mode)
//   flip / diffuse        phase flips and block-mean reflections
//                         (grover_classical.mojo:262-326)
//   unitary kernels       dense 2^m blocks (transformations_unitary.mojo:22-139):
//                         m = 4..6 on MFMA (FP64 / FP32), others on FMA
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <atomic>

#include "kernels.h"
#include "tile_ops.h"

namespace bfly {

namespace {

__device__ __forceinline__ uint64_t insert_zero64(uint64_t x, int pos) {
    const uint64_t lo = x & ((1ull << pos) - 1);
    return ((x >> pos) << (pos + 1)) | lo;
}


template <typename T>
__device__ __forceinline__ void load_params(const double* m, const float* mf, T* g);
template <>
__device__ __forceinline__ void load_params<double>(const double* m, const float*, double* g) {
#pragma unroll
    for (int i = 0; i < 8; ++i) g[i] = m[i];
}
template <>
__device__ __forceinline__ void load_params<float>(const double*, const float* mf, float* g) {
#pragma unroll
    for (int i = 0; i < 8; ++i) g[i] = mf[i];
}

// transform_h_scalar's (u +- v) * 0.7071067811865476 (transformations_scalar.mojo:241-283).
template <typename T>
__device__ __forceinline__ void pair_h(T& ur, T& ui, T& vr, T& vi) {
    const T s = static_cast<T>(0.7071067811865476);
    const T a = ur, b = ui, c = vr, d = vi;
    ur = (a + c) * s;
    ui = (b + d) * s;
    vr = (a - c) * s;
    vi = (b - d) * s;
}

// transform_ry_scalar (transformations_scalar.mojo:384-391): g[0]=cos, g[1]=sin.
template <typename T>
__device__ __forceinline__ void pair_ry(T& ur, T& ui, T& vr, T& vi, const T* g) {
    const T a = ur, b = ui, c = vr, d = vi;
    ur = a * g[0] - c * g[1];
    ui = b * g[0] - d * g[1];
    vr = a * g[1] + c * g[0];
    vi = b * g[1] + d * g[0];
}

template <typename T>
__device__ __forceinline__ void apply_pair_kind(int kind, T& ur, T& ui, T& vr, T& vi, const T* g) {
    switch (kind) {
        case I_H: pair_h(ur, ui, vr, vi); break;
        case I_X: { T t = ur; ur = vr; vr = t; t = ui; ui = vi; vi = t; } break;
        case I_RY: pair_ry(ur, ui, vr, vi, g); break;
        case I_DIAG: cmul(ur, ui, g[0], g[1]); cmul(vr, vi, g[2], g[3]); break;
        case I_PHASE1: cmul(vr, vi, g[2], g[3]); break;
        case I_NEG1: vr = -vr; vi = -vi; break;
        default: pair_gen(ur, ui, vr, vi, g); break;
    }
}

// ---------------------------------------------------------------- (a) gate
template <typename T>
__global__ __launch_bounds__(256) void gate_kernel(T* __restrict__ re, T* __restrict__ im, const GateLaunch G) {
    T g[8];
    load_params<T>(G.m, G.mf, g);
    const uint64_t tbit = 1ull << G.target;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; w < G.count; w += stride) {
        uint64_t i0 = w;
        for (int p = 0; p < G.npos; ++p) i0 = insert_zero64(i0, G.pos[p]);
        i0 |= G.cmask;
        if (G.kind == I_PHASE1 || G.kind == I_NEG1) {
            const uint64_t i1 = i0 | tbit;
            T vr = re[i1], vi = im[i1];
            if (G.kind == I_NEG1) { vr = -vr; vi = -vi; }
            else cmul(vr, vi, g[2], g[3]);
            re[i1] = vr;
            im[i1] = vi;
        } else {
            const uint64_t i1 = i0 | tbit;
            T ur = re[i0], ui = im[i0], vr = re[i1], vi = im[i1];
            apply_pair_kind<T>(G.kind, ur, ui, vr, vi, g);
            re[i0] = ur;
            im[i0] = ui;
            re[i1] = vr;
            im[i1] = vi;
        }
    }
}

// ---------------------------------------------------------------- (b) tile pass
template <typename T>
struct Vec;
template <>
struct Vec<double> { using type = double2; static constexpr int N = 2; };
template <>
struct Vec<float> { using type = float4; static constexpr int N = 4; };

// XOR swizzle of the LDS index: folds bits [S, 2S) into [0, S) so that the
// lanes of a sweep (which differ in the lowest non-slot bits) spread over
// the LDS banks.  S = 4 for 8-byte elements, 5 for 4-byte elements.
template <typename T>
__device__ __forceinline__ uint32_t swz(uint32_t j) {
    return lds_swizzle(j, sizeof(T));
}

template <int K>
__device__ __forceinline__ uint64_t local_to_global(uint32_t j, const TilePass& P) {
    uint64_t g = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) g |= static_cast<uint64_t>((j >> i) & 1u) << P.qpos[i];
    return g;
}

// Register copy of the per-sweep part of a TileInstr (parameters in the
// state's precision): 5 + 8 or 16 scalar registers.
template <typename T>
struct InstrRegs {
    uint32_t w0;       // kind | slot << 8 | cslots << 16 | dlocal << 24 (bytes of the record)
    uint32_t w1;       // obitpos | guard << 8
    uint32_t mm;       // mm1 | mm0 << 16
    uint32_t lmask, lval;
    T g[8];
};

template <typename T>
__device__ __forceinline__ InstrRegs<T> load_instr(const TileInstr* __restrict__ p) {
    InstrRegs<T> r;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
    r.w0 = w[0];
    r.w1 = w[1];
    r.mm = w[2];
    r.lmask = w[3];
    r.lval = w[4];
    load_params<T>(p->m, p->mf, r.g);
    return r;
}

__device__ __forceinline__ int sbyte(uint32_t w, int i) { return static_cast<int8_t>((w >> (8 * i)) & 0xffu); }
__device__ __forceinline__ uint32_t ubyte(uint32_t w, int i) { return (w >> (8 * i)) & 0xffu; }

// Sweeps of one tile pass on the tile staged in LDS (s_re, s_im): each of
// the 2^(K-S) groups of 2^S amplitudes spanned by the sweep's S slot bits
// (S = 3 or 4) is loaded into registers, every gate of the sweep runs in
// registers, and the group is written back.  Ends with a barrier after every
// sweep.
// Calls CALL with the compile-time slot index SL equal to `slot` (< S).
#define BFLY_SLOT_SWITCH(S, slot, CALL)                                    \
    switch (slot) {                                                        \
        case 0: { constexpr int SL = 0; CALL; } break;                     \
        case 1: { constexpr int SL = 1; CALL; } break;                     \
        case 2: { constexpr int SL = 2; if constexpr (SL < S) { CALL; } } break; \
        default: { constexpr int SL = 3; if constexpr (SL < S) { CALL; } } break; \
    }

template <typename T, int K, int S>
__device__ __forceinline__ void run_sweeps(T* s_re, T* s_im, const TilePass& P, const TileSweep* __restrict__ sweeps,
                                           const TileInstr* __restrict__ instrs, uint64_t base, uint32_t tid) {
    constexpr int G = 1 << S;
    constexpr uint32_t NG = (1u << K) / G;
    const int nsweeps = P.nsweeps;
    const int first_sweep = P.first_sweep;
    for (int s = 0; s < nsweeps; ++s) {
        const TileSweep sw = sweeps[first_sweep + s];
        if (tid < NG) {
            int srt[S];
#pragma unroll
            for (int a = 0; a < S; ++a) srt[a] = sw.s[a];
#pragma unroll
            for (int a = 0; a < S; ++a)
#pragma unroll
                for (int b2 = a + 1; b2 < S; ++b2)
                    if (srt[b2] < srt[a]) { const int t = srt[a]; srt[a] = srt[b2]; srt[b2] = t; }
            uint32_t jb = tid;
#pragma unroll
            for (int a = 0; a < S; ++a) jb = insert_zero32(jb, srt[a]);
            // member m of the group sits at local index jb | dep(m), dep(m) the
            // bits of m placed at the slots; swz(jb | dep(m)) = swz(jb) ^
            // swz(dep(m)) (swz is linear, the bits are disjoint), and the
            // planner stores the uniform second term in bytes (sw.doff), so
            // each member's address is one XOR
            constexpr int kLogT = sizeof(T) == 8 ? 3 : 2;
            const uint32_t jbs = swz<T>(jb) << kLogT;
            uint32_t off[G];
#pragma unroll
            for (int m = 0; m < G; ++m) off[m] = jbs ^ sw.doff[m];
            T ar[G], ai[G];
#pragma unroll
            for (int m = 0; m < G; ++m) {
                ar[m] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(s_re) + off[m]);
                ai[m] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(s_im) + off[m]);
            }
            const TileInstr* ip = instrs + sw.first;
            const int ni = sw.ninstr;
            for (int q = 0; q < ni; ++q) {
                const InstrRegs<T> in = load_instr<T>(ip + q);
                // the record's header and parameters arrive in one scalar-load
                // round trip (otherwise the parameter load is issued only after
                // the header tests, a second round trip per record)
#pragma unroll
                for (int e = 0; e < 8; ++e) asm volatile("" ::"s"(in.g[e]));
                if (ubyte(in.w1, 1) && (base & ip[q].omask) != ip[q].oval) continue;
                if ((jb & in.lmask) != in.lval) continue;
                const int kind = sbyte(in.w0, 0);
                const int slot = sbyte(in.w0, 1);
                const uint32_t cslots = ubyte(in.w0, 2);
                const int dlocal = sbyte(in.w0, 3);
                const T* g = in.g;
                if (S == 3 && kind == T_ROT3) {
                    // every slot, form 0: straight-line, so the register
                    // groups need no copies at control-flow joins
                    const T* h = sizeof(T) == 8 ? reinterpret_cast<const T*>(ip[q].m) : reinterpret_cast<const T*>(ip[q].mf);
                    slot_rot<T, G, 0>(ar, ai, h[0], false, h[1], h[2]);
                    slot_rot<T, G, 1>(ar, ai, h[3], false, h[4], h[5]);
                    if constexpr (S > 2) slot_rot<T, G, 2>(ar, ai, h[6], false, h[7], h[8]);
                } else if (kind == T_ROT) {
                    if (g[1] == T(0)) {
                        BFLY_SLOT_SWITCH(S, slot, (slot_rot<T, G, SL>(ar, ai, g[0], false, g[2], g[3])))
                    } else {
                        BFLY_SLOT_SWITCH(S, slot, (slot_rot<T, G, SL>(ar, ai, g[0], true, g[2], g[3])))
                    }
                } else if (S == 3 && kind == T_ROTS) {
                    const T* h = sizeof(T) == 8 ? reinterpret_cast<const T*>(ip[q].m) : reinterpret_cast<const T*>(ip[q].mf);
                    if (cslots & 1u) slot_rot<T, G, 0>(ar, ai, fabs_t(h[0]), signbit_t(h[0]), h[1], h[2]);
                    if (cslots & 2u) slot_rot<T, G, 1>(ar, ai, fabs_t(h[3]), signbit_t(h[3]), h[4], h[5]);
                    if constexpr (S > 2) {
                        if (cslots & 4u) slot_rot<T, G, 2>(ar, ai, fabs_t(h[6]), signbit_t(h[6]), h[7], h[8]);
                    }
                    if constexpr (S > 3) {
                        if (cslots & 8u) slot_rot<T, G, 3>(ar, ai, fabs_t(h[9]), signbit_t(h[9]), h[10], h[11]);
                    }
                } else if (kind == T_PH) {
                    BFLY_SLOT_SWITCH(S, slot, (slot_phase<T, G, SL>(ar, ai, g[0], g[1])))
                } else if (kind == T_SCALE) {
                    scale_all<T, G>(ar, ai, g[0], g[1]);
                } else if (kind == T_PAIR) {
                    BFLY_SLOT_SWITCH(S, slot, (slot_pairs<T, G, SL>(ar, ai, cslots, g)))
                } else {
                    uint32_t sel1 = __builtin_amdgcn_readfirstlane(in.mm & 0xffffu);
                    uint32_t sel0 = __builtin_amdgcn_readfirstlane(in.mm >> 16);
                    if (slot < 0) {
                        // target outside the slots: one bit per thread (local) or per tile (outer)
                        const int obitpos = sbyte(in.w1, 0);
                        const uint32_t bit = dlocal >= 0 ? ((jb >> dlocal) & 1u)
                                                         : static_cast<uint32_t>((base >> obitpos) & 1ull);
                        if (kind == T_NEG) {
                            neg_apply_bit<T, G>(ar, ai, sel1, sel0, bit);
                        } else {
                            diag_apply<T, G>(ar, ai, bit ? sel1 : 0u, bit ? 0u : sel0, g);
                        }
                    } else if (kind == T_NEG) {
                        if constexpr (S == 3) {
                            // member factors (+-1, exact) come with the record
#pragma unroll
                            for (int m = 0; m < G; ++m) {
                                ar[m] *= g[m];
                                ai[m] *= g[m];
                            }
                        } else {
                            neg_apply<T, G>(ar, ai, sel1);
                        }
                    } else {
                        diag_apply<T, G>(ar, ai, sel1, sel0, g);
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < G; ++m) {
                *reinterpret_cast<T*>(reinterpret_cast<char*>(s_re) + off[m]) = ar[m];
                *reinterpret_cast<T*>(reinterpret_cast<char*>(s_im) + off[m]) = ai[m];
            }
        }
        __syncthreads();
    }
}

// Threads of a tile block: one per group of 2^S amplitudes (at least a
// wave).  Minimum waves per SIMD the register budget must allow: the blocks
// that fit in 160 KiB of LDS (two at most), each spread over 4 SIMDs.
constexpr int tile_threads(int k, int s) { return k - s >= 6 ? (1 << (k - s)) : 64; }
constexpr int tile_waves_per_eu(int elem, int k, int s) {
    const int lds = 2 * elem << k;
    const int blocks = lds > 80 * 1024 ? 1 : 2;
    const int w = blocks * (tile_threads(k, s) / 64) / 4;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// The batched kernel keeps more per-tile state (item lookup): at most 4 waves
// per SIMD (2 for complex128) so that it never spills.
constexpr int batch_waves_per_eu(int elem, int k, int s) {
    const int cap = elem == 8 ? 2 : 4;
    return tile_waves_per_eu(elem, k, s) > cap ? cap : tile_waves_per_eu(elem, k, s);
}

template <int K>
__device__ __forceinline__ uint64_t tile_base(uint64_t t, const TilePass& P) {
    uint64_t b = 0;
    const int nouter = P.nouter;
    for (int i = 0; i < nouter; ++i) b |= ((t >> i) & 1ull) << P.opos[i];
    return b;
}

// (b) One tile pass.  A tile is 2^K amplitudes: local bit i is global bit
// P.qpos[i] (the low local bits are the low qubits, so VN consecutive local
// indices are VN consecutive amplitudes), the other bits come from the tile
// index.  NT = 2^(K-S) threads, one 2^S-amplitude group each per sweep.
// Blocks are persistent: block b handles tiles b, b + gridDim.x, ...; the
// next tile's HBM loads are issued into registers before the current tile's
// sweeps and land in LDS only after its stage-out, so HBM traffic overlaps
// the arithmetic.
template <typename T, int K, int S>
__global__ __launch_bounds__(tile_threads(K, S), tile_waves_per_eu(sizeof(T), K, S)) void tile_pass_kernel(
        T* __restrict__ re, T* __restrict__ im, const TilePass P, const TileSweep* __restrict__ sweeps,
        const TileInstr* __restrict__ instrs) {
    constexpr uint32_t TS = 1u << K;
    constexpr uint32_t NT = tile_threads(K, S);  // threads per block
    // register prefetch of the next tile only where two waves per SIMD are
    // the occupancy anyway; otherwise the co-resident blocks hide the loads
    constexpr bool kPrefetch = tile_waves_per_eu(sizeof(T), K, S) <= 2;
    constexpr int VN = Vec<T>::N;
    constexpr uint32_t NV = TS / VN;               // 16-byte vectors per tile (re or im)
    constexpr uint32_t PV = (NV + NT - 1) / NT;    // vectors per thread
    using V = typename Vec<T>::type;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* s_re = reinterpret_cast<T*>(lds_raw);
    T* s_im = s_re + TS;
    const uint32_t tid = threadIdx.x;
    const uint64_t ntiles = 1ull << (P.n - P.k);
    uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;

    // local index of this thread's c-th vector is (tid + c*NT)*VN; its global
    // offset splits into a per-thread part and a uniform part (disjoint bits)
    const uint64_t off_t = local_to_global<K>(tid * VN, P);
    V pr[PV], pi[PV];
    auto load_tile = [&](uint64_t b) {
#pragma unroll
        for (uint32_t c = 0; c < PV; ++c) {
            if (tid + c * NT < NV) {
                const uint64_t g = b + off_t + local_to_global<K>(c * NT * VN, P);
                pr[c] = *reinterpret_cast<const V*>(re + g);
                pi[c] = *reinterpret_cast<const V*>(im + g);
            }
        }
    };
    uint64_t base = tile_base<K>(tile, P);
    load_tile(base);

    for (;;) {
        // registers -> LDS
#pragma unroll
        for (uint32_t c = 0; c < PV; ++c) {
            if (tid + c * NT < NV) {
                const uint32_t j = (tid + c * NT) * VN;
                const T* a = reinterpret_cast<const T*>(&pr[c]);
                const T* b = reinterpret_cast<const T*>(&pi[c]);
#pragma unroll
                for (int v = 0; v < VN; ++v) {
                    s_re[swz<T>(j + v)] = a[v];
                    s_im[swz<T>(j + v)] = b[v];
                }
            }
        }
        __syncthreads();
        const uint64_t next = tile + gridDim.x;
        uint64_t nbase = 0;
        if (next < ntiles) {
            nbase = tile_base<K>(next, P);
            if (kPrefetch) load_tile(nbase);  // in flight during the sweeps
        }

        run_sweeps<T, K, S>(s_re, s_im, P, sweeps, instrs, base, tid);

        // stage out: LDS -> HBM with 16-byte stores
#pragma unroll
        for (uint32_t c = 0; c < PV; ++c) {
            if (tid + c * NT < NV) {
                const uint32_t j = (tid + c * NT) * VN;
                V vr, vi;
                T* a = reinterpret_cast<T*>(&vr);
                T* b = reinterpret_cast<T*>(&vi);
#pragma unroll
                for (int v = 0; v < VN; ++v) {
                    a[v] = s_re[swz<T>(j + v)];
                    b[v] = s_im[swz<T>(j + v)];
                }
                const uint64_t g = base + off_t + local_to_global<K>(c * NT * VN, P);
                *reinterpret_cast<V*>(re + g) = vr;
                *reinterpret_cast<V*>(im + g) = vi;
            }
        }
        if (next >= ntiles) break;
        __syncthreads();  // every stage-out read is done before the next tile lands in LDS
        tile = next;
        base = nbase;
        if (!kPrefetch) load_tile(base);
    }
}

// Batched tile passes: one launch runs the next tile pass of many circuits
// (each with its own state and plan).  Item i owns global tiles
// [items[i].tile_begin, items[i].tile_begin + 2^(n_i - K)); blocks walk the
// global tile range persistently and find their item by binary search.
template <typename T, int K, int S>
__global__ __launch_bounds__(tile_threads(K, S), batch_waves_per_eu(sizeof(T), K, S)) void tile_batch_kernel(
        const BatchItem* __restrict__ items, int nitems, uint64_t total_tiles) {
    constexpr uint32_t TS = 1u << K;
    constexpr uint32_t NT = tile_threads(K, S);
    constexpr int VN = Vec<T>::N;
    constexpr uint32_t NV = TS / VN;
    using V = typename Vec<T>::type;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* s_re = reinterpret_cast<T*>(lds_raw);
    T* s_im = s_re + TS;
    const uint32_t tid = threadIdx.x;
    for (uint64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
        int lo = 0, hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].tile_begin <= t) lo = mid;
            else hi = mid - 1;
        }
        const BatchItem& it = items[lo];
        T* re = static_cast<T*>(it.re);
        T* im = static_cast<T*>(it.im);
        const uint64_t base = tile_base<K>(t - it.tile_begin, it.P);
        for (uint32_t e = tid; e < NV; e += NT) {
            const uint32_t j = e * VN;
            const uint64_t g = base + local_to_global<K>(j, it.P);
            const V vr = *reinterpret_cast<const V*>(re + g);
            const V vi = *reinterpret_cast<const V*>(im + g);
            const T* a = reinterpret_cast<const T*>(&vr);
            const T* b = reinterpret_cast<const T*>(&vi);
#pragma unroll
            for (int v = 0; v < VN; ++v) {
                s_re[swz<T>(j + v)] = a[v];
                s_im[swz<T>(j + v)] = b[v];
            }
        }
        __syncthreads();
        run_sweeps<T, K, S>(s_re, s_im, it.P, it.sweeps, it.instrs, base, tid);
        for (uint32_t e = tid; e < NV; e += NT) {
            const uint32_t j = e * VN;
            V vr, vi;
            T* a = reinterpret_cast<T*>(&vr);
            T* b = reinterpret_cast<T*>(&vi);
#pragma unroll
            for (int v = 0; v < VN; ++v) {
                a[v] = s_re[swz<T>(j + v)];
                b[v] = s_im[swz<T>(j + v)];
            }
            const uint64_t g = base + local_to_global<K>(j, it.P);
            *reinterpret_cast<V*>(re + g) = vr;
            *reinterpret_cast<V*>(im + g) = vi;
        }
        __syncthreads();  // stage-out reads finish before the next tile lands in LDS
    }
}


// Kernel attributes (dynamic LDS beyond 64 KiB) and the persistent grid size
// (blocks that fit on the device at once) for one kernel instance, cached per
// device.  The cache entries are atomics: threads launching on separate
// handles may race to fill an entry and compute the same value, which is
// harmless; a grid sized for one device is never reused on another.
constexpr int kGridCacheDevices = 64;
struct GridCache {
    std::atomic<int> cap[kGridCacheDevices] = {};
};

template <typename KernelT>
hipError_t persistent_grid(KernelT kernel, int nt, size_t lds, GridCache& cache, int* grid_cap) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kGridCacheDevices) return hipErrorInvalidDevice;
    const int cached = cache.cap[dev].load(std::memory_order_acquire);
    if (cached > 0) {
        *grid_cap = cached;
        return hipSuccess;
    }
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
    if (e != hipSuccess) return e;
    int per_cu = 0, cus = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, lds);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const int cap = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
    cache.cap[dev].store(cap, std::memory_order_release);
    *grid_cap = cap;
    return hipSuccess;
}

// Streaming launches use one block per unit of work (a tile, a tile pair, a
// vector group): the dispatcher starts a new block whenever one finishes, so
// the loads of some blocks overlap the arithmetic and stores of others on
// every CU.  Measured on MI355X: an in-place read + write of 2 x 8 GiB runs
// at 5.77 TB/s with one 16-byte vector per thread and a block per 256
// vectors, 4.5-5.4 TB/s with persistent grids of 1024-65536 blocks; the
// 30-qubit layered circuit's tile passes run 11 % faster.  BFLY_GRID=
// persistent restores the grids of the blocks the device holds at once
// (measurement).
bool persistent_grids() {
    static const bool p = [] {
        const char* g = std::getenv("BFLY_GRID");
        return g && g[0] == 'p';
    }();
    return p;
}
// `threads` per block: a dispatch covers at most 2^32 - 1 work-items per
// dimension, so larger grids (from 2^33 amplitudes) are capped and the
// kernels' grid-stride loops cover the rest.
inline unsigned full_grid(uint64_t units, int cap, unsigned threads) {
    if (persistent_grids()) return static_cast<unsigned>(units < static_cast<uint64_t>(cap) ? units : cap);
    const uint64_t max_blocks = 0xffffffffull / threads;
    return static_cast<unsigned>(units < max_blocks ? units : max_blocks);
}

template <typename T, int K, int S>
hipError_t launch_tile_k(T* re, T* im, const TilePass& P, const TileSweep* sw, const TileInstr* in, hipStream_t st) {
    constexpr int NT = tile_threads(K, S);
    const size_t lds = 2ull * (1ull << K) * sizeof(T);
    static GridCache grid_cache;
    int grid_cap = 0;
    hipError_t e = persistent_grid(tile_pass_kernel<T, K, S>, NT, lds, grid_cache, &grid_cap);
    if (e != hipSuccess) return e;
    if (!re) return hipSuccess;  // preparation only
    const uint64_t ntiles = 1ull << (P.n - P.k);
    const unsigned grid = full_grid(ntiles, grid_cap, NT);
    hipLaunchKernelGGL((tile_pass_kernel<T, K, S>), dim3(grid), dim3(NT), lds, st, re, im, P, sw, in);
    return hipGetLastError();
}

template <typename T, int K, int S>
hipError_t launch_batch_k(const BatchItem* items, int nitems, uint64_t total_tiles, hipStream_t st) {
    constexpr int NT = tile_threads(K, S);
    const size_t lds = 2ull * (1ull << K) * sizeof(T);
    static GridCache grid_cache;
    int grid_cap = 0;
    hipError_t e = persistent_grid(tile_batch_kernel<T, K, S>, NT, lds, grid_cache, &grid_cap);
    if (e != hipSuccess) return e;
    if (!items) return hipSuccess;  // preparation only
    const unsigned grid = full_grid(total_tiles, grid_cap, NT);
    hipLaunchKernelGGL((tile_batch_kernel<T, K, S>), dim3(grid), dim3(NT), lds, st, items, nitems, total_tiles);
    return hipGetLastError();
}

// Sweep widths compiled: 4 slots for every tile size, 3 slots for tiles of
// 2^10 .. 2^13 amplitudes.
template <typename T, int S>
hipError_t launch_batch_s(int k, const BatchItem* items, int nitems, uint64_t total_tiles, hipStream_t st) {
    switch (k) {
        case 4: if constexpr (S == 4) return launch_batch_k<T, 4, S>(items, nitems, total_tiles, st); break;
        case 5: if constexpr (S == 4) return launch_batch_k<T, 5, S>(items, nitems, total_tiles, st); break;
        case 6: if constexpr (S == 4) return launch_batch_k<T, 6, S>(items, nitems, total_tiles, st); break;
        case 7: if constexpr (S == 4) return launch_batch_k<T, 7, S>(items, nitems, total_tiles, st); break;
        case 8: if constexpr (S == 4) return launch_batch_k<T, 8, S>(items, nitems, total_tiles, st); break;
        case 9: if constexpr (S == 4) return launch_batch_k<T, 9, S>(items, nitems, total_tiles, st); break;
        case 10: return launch_batch_k<T, 10, S>(items, nitems, total_tiles, st);
        case 11: return launch_batch_k<T, 11, S>(items, nitems, total_tiles, st);
        case 12: return launch_batch_k<T, 12, S>(items, nitems, total_tiles, st);
        case 13: return launch_batch_k<T, 13, S>(items, nitems, total_tiles, st);
    }
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_batch_t(int k, int slots, const BatchItem* items, int nitems, uint64_t total_tiles, hipStream_t st) {
    if (slots == 3) return launch_batch_s<T, 3>(k, items, nitems, total_tiles, st);
    if (slots == 4) return launch_batch_s<T, 4>(k, items, nitems, total_tiles, st);
    return hipErrorInvalidValue;
}

template <typename T, int S>
hipError_t launch_tile_s(T* re, T* im, const TilePass& P, const TileSweep* sw, const TileInstr* in, hipStream_t st) {
    switch (P.k) {
        case 4: if constexpr (S == 4) return launch_tile_k<T, 4, S>(re, im, P, sw, in, st); break;
        case 5: if constexpr (S == 4) return launch_tile_k<T, 5, S>(re, im, P, sw, in, st); break;
        case 6: if constexpr (S == 4) return launch_tile_k<T, 6, S>(re, im, P, sw, in, st); break;
        case 7: if constexpr (S == 4) return launch_tile_k<T, 7, S>(re, im, P, sw, in, st); break;
        case 8: if constexpr (S == 4) return launch_tile_k<T, 8, S>(re, im, P, sw, in, st); break;
        case 9: if constexpr (S == 4) return launch_tile_k<T, 9, S>(re, im, P, sw, in, st); break;
        case 10: return launch_tile_k<T, 10, S>(re, im, P, sw, in, st);
        case 11: return launch_tile_k<T, 11, S>(re, im, P, sw, in, st);
        case 12: return launch_tile_k<T, 12, S>(re, im, P, sw, in, st);
        case 13: return launch_tile_k<T, 13, S>(re, im, P, sw, in, st);
    }
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_tile_t(T* re, T* im, const TilePass& P, const TileSweep* sw, const TileInstr* in, hipStream_t st) {
    if (P.slots == 3) return launch_tile_s<T, 3>(re, im, P, sw, in, st);
    if (P.slots == 4) return launch_tile_s<T, 4>(re, im, P, sw, in, st);
    return hipErrorInvalidValue;
}

// ------------------------------------------------ (b'') dense blocks on MFMA
// BLOCK mode (complex64): a tile of 2^K amplitudes (K = 9 .. 12) spanned by
// the block's 5 slots and K - 5 column bits is staged in LDS.  Each wave takes
// 16 columns at a time: the 32 amplitudes of a column, multiplied by their
// pre-phase, form the B operand (64 x 16: re over im) of
//   [out_re; out_im] = [[Ur, -Ui], [Ui, Ur]] [in_re; in_im]
// computed by 64 v_mfma_f32_16x16x4f32 (FP32 products, FP32 accumulation);
// the 64 x 64 A operand stays in registers for the whole pass.
// Phase records contribute exp(i (c0 + sum_j c_j bit_j)) per column: the
// angles c are summed in FP64 (exact for dyadic pi / 2^k ladders) once per
// tile and column, turned into factors by FP64 sincos, kept in LDS, and each
// lane forms its 8 element factors by FP64 products rounded to FP32 once.
// The next tile's amplitudes are loaded into registers while the current
// tile is multiplied.
typedef float f4_t __attribute__((ext_vector_type(4)));

struct Z64 {
    double r, i;
};
__device__ __forceinline__ Z64 zmul(Z64 a, Z64 b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ Z64 zexp(double t) {
    if (t == 0.0) return {1.0, 0.0};
    double sn, cs;
    sincos(t, &sn, &cs);
    return {cs, sn};
}

// The 8 factors of elements e = base_bits + {0, s1, s2, s1 + s2} + {0, s3}
// (four bit positions of e): f[m + 4 h] for m over (s1, s2), h over s3.
__device__ __forceinline__ void element_factors(const Z64 z0, const Z64 z1, const Z64 z2, const Z64 z3, float* fr,
                                                float* fi) {
    Z64 f[8];
    f[0] = z0;
    f[1] = zmul(f[0], z1);
    f[2] = zmul(f[0], z2);
    f[3] = zmul(f[1], z2);
#pragma unroll
    for (int m = 0; m < 4; ++m) f[4 + m] = zmul(f[m], z3);
#pragma unroll
    for (int m = 0; m < 8; ++m) { fr[m] = static_cast<float>(f[m].r); fi[m] = static_cast<float>(f[m].i); }
}

// z_j = (column factor j) x (tile factor j), FP64.
__device__ __forceinline__ void column_factors(const double2* cf, const double2* tf, Z64* z) {
#pragma unroll
    for (int j = 0; j <= kBlockQubits; ++j) z[j] = zmul(Z64{cf[j].x, cf[j].y}, Z64{tf[j].x, tf[j].y});
}

template <int K>
__global__ __launch_bounds__(256) void block_mfma_kernel(float* __restrict__ re, float* __restrict__ im,
                                                         const TilePass P, const BlockPass B,
                                                         const float* __restrict__ mat,
                                                         const double* __restrict__ tab,
                                                         const double* __restrict__ norm2) {
    constexpr uint32_t TS = 1u << K;
    constexpr int NT = 256;
    constexpr int VN = 4;
    constexpr uint32_t NV = TS / VN;
    constexpr int NPF = static_cast<int>((NV + NT - 1) / NT);   // float4 per thread and part
    constexpr int NCOL = static_cast<int>(TS / kBlockDim);       // 32-amplitude columns per tile
    constexpr int NGROUP = NCOL / 16;                            // 16-column groups per tile
    constexpr int NZ = kBlockQubits + 1;
    static_assert(NT * VN >= 1024 || NPF == 1, "part stride");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* s_re = reinterpret_cast<float*>(lds_raw);
    float* s_im = s_re + TS;
    double2* s_cf = reinterpret_cast<double2*>(s_im + TS);       // [2][NCOL][NZ]: pre, post column factors
    double2* s_tf = s_cf + 2 * NCOL * NZ;                        // [2 buffers][2][NZ]: tile factors
    const uint32_t tid = threadIdx.x;
    const int lane = static_cast<int>(tid & 63u);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
    const int g = lane >> 4;  // k offset of the lane's B / A fragments, row quad of its output

    float afrag[4][16];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int s2 = 0; s2 < 16; ++s2) afrag[r][s2] = mat[(16 * r + (lane & 15)) * 64 + 4 * s2 + g];

    uint32_t qmask = 0;
#pragma unroll
    for (int j = 0; j < kBlockQubits; ++j) qmask |= 1u << B.qloc[j];
    const uint32_t wmask = (TS - 1) & ~qmask;
    auto eloc = [&](int e) {
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < kBlockQubits; ++j) x |= static_cast<uint32_t>((e >> j) & 1) << B.qloc[j];
        return x;
    };
    auto col_local = [&](uint32_t col) {   // deposit of the column bits into wmask
        uint32_t colloc = 0, x = col, m = wmask;
        while (m) {
            const uint32_t low = m & (~m + 1);
            if (x & 1) colloc |= low;
            x >>= 1;
            m ^= low;
        }
        return colloc;
    };
    // Deposits and the LDS swizzle are linear over GF(2), so every address
    // splits into a per-lane part computed here and a uniform part.
    const uint32_t lane_j = VN * tid;                       // first local index of the thread's float4 (part 0)
    const uint64_t off_lane = local_to_global<K>(lane_j & (TS - 1), P);
    const uint32_t lds_lane = swz<float>(lane_j & (TS - 1));
    const uint32_t dl = col_local(static_cast<uint32_t>(lane & 15));
    const uint32_t rl_in = swz<float>(dl ^ eloc(g));        // B-fragment rows e = g + 4 m
    const uint32_t rl_out = swz<float>(dl ^ eloc(4 * g));   // output rows e = 16 h + 4 g + v
    const double scale = B.renorm ? 1.0 / sqrt(norm2[0]) : 1.0;
    const bool pre_any = B.pre_tab != 0, post_any = B.post_tab != 0;
    const double* outer_ang = tab + B.tab_off + 2 * NCOL * NZ * 2;   // [2][nouter][NZ]

    // column factors stay in LDS for the whole launch
    if (pre_any || post_any) {
        const double2* cf = reinterpret_cast<const double2*>(tab + B.tab_off);
        for (uint32_t i = tid; i < 2u * NCOL * NZ; i += NT) s_cf[i] = cf[i];
    }
    // the tile's factors exp(i sum of its outer-bit angles), pre scaled by `scale`
    auto tile_factors = [&](uint64_t t, int buf) {
        if (tid < 2 * NZ) {
            const int w = static_cast<int>(tid) / NZ, j = static_cast<int>(tid) % NZ;
            const double* a = outer_ang + static_cast<size_t>(w) * P.nouter * NZ;
            double x = 0.0;
            for (int i = 0; i < P.nouter; ++i)
                if ((t >> i) & 1ull) x += a[i * NZ + j];
            Z64 z = zexp(x);
            if (w == 0 && j == 0) { z.r *= scale; z.i *= scale; }
            s_tf[buf * 2 * NZ + tid] = make_double2(z.r, z.i);
        }
    };

    const uint64_t ntiles = 1ull << (P.n - P.k);
    float4 pa[NPF], pb[NPF];
    auto fetch = [&](uint64_t base) {
#pragma unroll
        for (int c = 0; c < NPF; ++c) {
            if (NPF > 1 || tid < NV) {
                const uint64_t gi = base + off_lane + local_to_global<K>(static_cast<uint32_t>(c * NT * VN), P);
                pa[c] = *reinterpret_cast<const float4*>(re + gi);
                pb[c] = *reinterpret_cast<const float4*>(im + gi);
            }
        }
    };
    uint64_t tile = blockIdx.x;
    if (tile < ntiles) {
        fetch(tile_base<K>(tile, P));
        if (pre_any || post_any) tile_factors(tile, 0);
    }
    for (int buf = 0; tile < ntiles; tile += gridDim.x, buf ^= 1) {
        const uint64_t base = tile_base<K>(tile, P);
#pragma unroll
        for (int c = 0; c < NPF; ++c) {
            if (NPF > 1 || tid < NV) {
                const uint32_t a = lds_lane + static_cast<uint32_t>(c * NT * VN);
                s_re[a] = pa[c].x; s_re[a ^ 1u] = pa[c].y; s_re[a ^ 2u] = pa[c].z; s_re[a ^ 3u] = pa[c].w;
                s_im[a] = pb[c].x; s_im[a ^ 1u] = pb[c].y; s_im[a ^ 2u] = pb[c].z; s_im[a ^ 3u] = pb[c].w;
            }
        }
        __syncthreads();
        {
            const uint64_t nt = tile + gridDim.x;
            if (nt < ntiles) {
                fetch(tile_base<K>(nt, P));
                if (pre_any || post_any) tile_factors(nt, buf ^ 1);
            }
        }
        const double2* tf = s_tf + buf * 2 * NZ;
        for (int cg = wave; cg < NGROUP; cg += NT / 64) {
            const uint32_t col = static_cast<uint32_t>(cg * 16 + (lane & 15));
            const uint32_t ucg = swz<float>(col_local(static_cast<uint32_t>(cg * 16)));
            float bfrag[16];
            if (pre_any) {
                // pre-phase factors of the lane's 8 elements e = g + 4 m
                float fr[8], fi[8];
                Z64 z[NZ];
                column_factors(s_cf + col * NZ, tf, z);
                Z64 b0 = z[0];
                if (g & 1) b0 = zmul(b0, z[1]);
                if (g & 2) b0 = zmul(b0, z[2]);
                element_factors(b0, z[3], z[4], z[5], fr, fi);
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const uint32_t j = rl_in ^ ucg ^ swz<float>(eloc(4 * m));
                    const float xr = s_re[j], xi = s_im[j];
                    bfrag[m] = xr * fr[m] - xi * fi[m];
                    bfrag[m + 8] = xr * fi[m] + xi * fr[m];
                }
            } else {
                const float sc = static_cast<float>(scale);
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const uint32_t j = rl_in ^ ucg ^ swz<float>(eloc(4 * m));
                    bfrag[m] = s_re[j] * sc;
                    bfrag[m + 8] = s_im[j] * sc;
                }
            }
            f4_t acc[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[r] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s2 = 0; s2 < 16; ++s2)
                    acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag[r][s2], bfrag[s2], acc[r], 0, 0, 0);
            }
            // lane holds rows 16 r + 4 g + v: re of e = 16 h + 4 g + v (r = h), im (r = 2 + h)
            if (post_any) {
                float pr[8], pi2[8];
                Z64 z[NZ];
                column_factors(s_cf + NCOL * NZ + col * NZ, tf + NZ, z);
                Z64 b0 = z[0];
                if (g & 1) b0 = zmul(b0, z[3]);
                if (g & 2) b0 = zmul(b0, z[4]);
                element_factors(b0, z[1], z[2], z[5], pr, pi2);
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float yr = acc[h][v], yi = acc[2 + h][v];
                        const int q = 4 * h + v;
                        const uint32_t j = rl_out ^ ucg ^ swz<float>(eloc(16 * h + v));
                        s_re[j] = yr * pr[q] - yi * pi2[q];
                        s_im[j] = yr * pi2[q] + yi * pr[q];
                    }
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const uint32_t j = rl_out ^ ucg ^ swz<float>(eloc(16 * h + v));
                        s_re[j] = acc[h][v];
                        s_im[j] = acc[2 + h][v];
                    }
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NPF; ++c) {
            if (NPF > 1 || tid < NV) {
                const uint32_t a = lds_lane + static_cast<uint32_t>(c * NT * VN);
                float4 x, y;
                x.x = s_re[a]; x.y = s_re[a ^ 1u]; x.z = s_re[a ^ 2u]; x.w = s_re[a ^ 3u];
                y.x = s_im[a]; y.y = s_im[a ^ 1u]; y.z = s_im[a ^ 2u]; y.w = s_im[a ^ 3u];
                const uint64_t gi = base + off_lane + local_to_global<K>(static_cast<uint32_t>(c * NT * VN), P);
                *reinterpret_cast<float4*>(re + gi) = x;
                *reinterpret_cast<float4*>(im + gi) = y;
            }
        }
        __syncthreads();
    }
}

// complex128 BLOCK mode on v_mfma_f64_16x16x4f64.  Only the upper half of the
// real form, A = [Ur, -Ui] (32 x 64), stays in registers: the real outputs are
// A [re; im] and the imaginary outputs A [im; -re].  The f64 instruction's
// result layout differs from the f32 one: lane l holds rows (l / 16) + 4 v
// of its 16-row block (v = 0..3), column l % 16.  Tiles of 2^K amplitudes
// (K = 9 .. 11), 16-byte vectors of two amplitudes, phase factors in FP64.
typedef double d4_t __attribute__((ext_vector_type(4)));

template <int K>
__global__ __launch_bounds__(256) void block_mfma64_kernel(double* __restrict__ re, double* __restrict__ im,
                                                           const TilePass P, const BlockPass B,
                                                           const double* __restrict__ mat,
                                                           const double* __restrict__ tab,
                                                           const double* __restrict__ norm2) {
    constexpr uint32_t TS = 1u << K;
    constexpr int NT = 256;
    constexpr int VN = 2;
    constexpr uint32_t NV = TS / VN;
    constexpr int NPF = static_cast<int>((NV + NT - 1) / NT);
    constexpr int NCOL = static_cast<int>(TS / kBlockDim);
    constexpr int NGROUP = NCOL / 16;
    constexpr int NZ = kBlockQubits + 1;
    static_assert(NT * VN >= 512 || NPF == 1, "part stride");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double* s_re = reinterpret_cast<double*>(lds_raw);
    double* s_im = s_re + TS;
    double2* s_cf = reinterpret_cast<double2*>(s_im + TS);
    double2* s_tf = s_cf + 2 * NCOL * NZ;
    const uint32_t tid = threadIdx.x;
    const int lane = static_cast<int>(tid & 63u);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
    const int g = lane >> 4;

    // A fragments: row block r (rows 16 r + lane % 16), k step s2 (k = 4 s2 + g)
    double afrag[2][16];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int s2 = 0; s2 < 16; ++s2) afrag[r][s2] = mat[(16 * r + (lane & 15)) * 64 + 4 * s2 + g];

    uint32_t qmask = 0;
#pragma unroll
    for (int j = 0; j < kBlockQubits; ++j) qmask |= 1u << B.qloc[j];
    const uint32_t wmask = (TS - 1) & ~qmask;
    auto eloc = [&](int e) {
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < kBlockQubits; ++j) x |= static_cast<uint32_t>((e >> j) & 1) << B.qloc[j];
        return x;
    };
    auto col_local = [&](uint32_t col) {
        uint32_t colloc = 0, x = col, m = wmask;
        while (m) {
            const uint32_t low = m & (~m + 1);
            if (x & 1) colloc |= low;
            x >>= 1;
            m ^= low;
        }
        return colloc;
    };
    const uint32_t lane_j = VN * tid;
    const uint64_t off_lane = local_to_global<K>(lane_j & (TS - 1), P);
    const uint32_t lds_lane = swz<double>(lane_j & (TS - 1));
    const uint32_t dl = col_local(static_cast<uint32_t>(lane & 15));
    const uint32_t rl_in = swz<double>(dl ^ eloc(g));        // B rows e = g + 4 m
    const uint32_t rl_out = swz<double>(dl ^ eloc(g));       // output rows e = g + 4 v + 16 h
    const double scale = B.renorm ? 1.0 / sqrt(norm2[0]) : 1.0;
    const bool pre_any = B.pre_tab != 0, post_any = B.post_tab != 0;
    const double* outer_ang = tab + B.tab_off + 2 * NCOL * NZ * 2;

    if (pre_any || post_any) {
        const double2* cf = reinterpret_cast<const double2*>(tab + B.tab_off);
        for (uint32_t i = tid; i < 2u * NCOL * NZ; i += NT) s_cf[i] = cf[i];
    }
    auto tile_factors = [&](uint64_t t, int buf) {
        if (tid < 2 * NZ) {
            const int w = static_cast<int>(tid) / NZ, j = static_cast<int>(tid) % NZ;
            const double* a = outer_ang + static_cast<size_t>(w) * P.nouter * NZ;
            double x = 0.0;
            for (int i = 0; i < P.nouter; ++i)
                if ((t >> i) & 1ull) x += a[i * NZ + j];
            Z64 z = zexp(x);
            if (w == 0 && j == 0) { z.r *= scale; z.i *= scale; }
            s_tf[buf * 2 * NZ + tid] = make_double2(z.r, z.i);
        }
    };

    const uint64_t ntiles = 1ull << (P.n - P.k);
    double2 pa[NPF], pb[NPF];
    auto fetch = [&](uint64_t base) {
#pragma unroll
        for (int c = 0; c < NPF; ++c) {
            if (NPF > 1 || tid < NV) {
                const uint64_t gi = base + off_lane + local_to_global<K>(static_cast<uint32_t>(c * NT * VN), P);
                pa[c] = *reinterpret_cast<const double2*>(re + gi);
                pb[c] = *reinterpret_cast<const double2*>(im + gi);
            }
        }
    };
    uint64_t tile = blockIdx.x;
    if (tile < ntiles) {
        fetch(tile_base<K>(tile, P));
        if (pre_any || post_any) tile_factors(tile, 0);
    }
    for (int buf = 0; tile < ntiles; tile += gridDim.x, buf ^= 1) {
        const uint64_t base = tile_base<K>(tile, P);
#pragma unroll
        for (int c = 0; c < NPF; ++c) {
            if (NPF > 1 || tid < NV) {
                const uint32_t a = lds_lane + static_cast<uint32_t>(c * NT * VN);
                s_re[a] = pa[c].x; s_re[a ^ 1u] = pa[c].y;
                s_im[a] = pb[c].x; s_im[a ^ 1u] = pb[c].y;
            }
        }
        __syncthreads();
        {
            const uint64_t nt = tile + gridDim.x;
            if (nt < ntiles) {
                fetch(tile_base<K>(nt, P));
                if (pre_any || post_any) tile_factors(nt, buf ^ 1);
            }
        }
        const double2* tf = s_tf + buf * 2 * NZ;
        for (int cg = wave; cg < NGROUP; cg += NT / 64) {
            const uint32_t col = static_cast<uint32_t>(cg * 16 + (lane & 15));
            const uint32_t ucg = swz<double>(col_local(static_cast<uint32_t>(cg * 16)));
            double br[8], bi[8];   // the lane's elements e = g + 4 m, pre-phased
            if (pre_any) {
                Z64 z[NZ];
                column_factors(s_cf + col * NZ, tf, z);
                Z64 b0 = z[0];
                if (g & 1) b0 = zmul(b0, z[1]);
                if (g & 2) b0 = zmul(b0, z[2]);
                Z64 f[8];
                f[0] = b0;
                f[1] = zmul(f[0], z[3]);
                f[2] = zmul(f[0], z[4]);
                f[3] = zmul(f[1], z[4]);
#pragma unroll
                for (int m = 0; m < 4; ++m) f[4 + m] = zmul(f[m], z[5]);
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const uint32_t j = rl_in ^ ucg ^ swz<double>(eloc(4 * m));
                    const double xr = s_re[j], xi = s_im[j];
                    br[m] = xr * f[m].r - xi * f[m].i;
                    bi[m] = xr * f[m].i + xi * f[m].r;
                }
            } else {
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const uint32_t j = rl_in ^ ucg ^ swz<double>(eloc(4 * m));
                    br[m] = s_re[j] * scale;
                    bi[m] = s_im[j] * scale;
                }
            }
            // [re'] = A [re; im], [im'] = A [im; -re]
            d4_t acc_r[2], acc_i[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                acc_r[r] = d4_t{0.0, 0.0, 0.0, 0.0};
                acc_i[r] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s2 = 0; s2 < 8; ++s2) {
                    acc_r[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][s2], br[s2], acc_r[r], 0, 0, 0);
                    acc_i[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][s2], bi[s2], acc_i[r], 0, 0, 0);
                }
#pragma unroll
                for (int s2 = 0; s2 < 8; ++s2) {
                    acc_r[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][8 + s2], bi[s2], acc_r[r], 0, 0, 0);
                    acc_i[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][8 + s2], -br[s2], acc_i[r], 0, 0, 0);
                }
            }
            // lane holds e = 16 h + g + 4 v (row block h, accumulator v)
            if (post_any) {
                Z64 z[NZ];
                column_factors(s_cf + NCOL * NZ + col * NZ, tf + NZ, z);
                Z64 b0 = z[0];
                if (g & 1) b0 = zmul(b0, z[1]);
                if (g & 2) b0 = zmul(b0, z[2]);
                Z64 f[8];   // index v + 4 h: v bits are e bits 2, 3; h is e bit 4
                f[0] = b0;
                f[1] = zmul(f[0], z[3]);
                f[2] = zmul(f[0], z[4]);
                f[3] = zmul(f[1], z[4]);
#pragma unroll
                for (int v = 0; v < 4; ++v) f[4 + v] = zmul(f[v], z[5]);
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double yr = acc_r[h][v], yi = acc_i[h][v];
                        const Z64 fq = f[v + 4 * h];
                        const uint32_t j = rl_out ^ ucg ^ swz<double>(eloc(16 * h + 4 * v));
                        s_re[j] = yr * fq.r - yi * fq.i;
                        s_im[j] = yr * fq.i + yi * fq.r;
                    }
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const uint32_t j = rl_out ^ ucg ^ swz<double>(eloc(16 * h + 4 * v));
                        s_re[j] = acc_r[h][v];
                        s_im[j] = acc_i[h][v];
                    }
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NPF; ++c) {
            if (NPF > 1 || tid < NV) {
                const uint32_t a = lds_lane + static_cast<uint32_t>(c * NT * VN);
                double2 x, y;
                x.x = s_re[a]; x.y = s_re[a ^ 1u];
                y.x = s_im[a]; y.y = s_im[a ^ 1u];
                const uint64_t gi = base + off_lane + local_to_global<K>(static_cast<uint32_t>(c * NT * VN), P);
                *reinterpret_cast<double2*>(re + gi) = x;
                *reinterpret_cast<double2*>(im + gi) = y;
            }
        }
        __syncthreads();
    }
}

template <int K>
hipError_t launch_block64_k(double* re, double* im, const TilePass& P, const BlockPass& B, const double* mat,
                            const double* tab, const double* norm2, hipStream_t st) {
    const size_t lds = 2ull * (1ull << K) * sizeof(double) + 2ull * ((1ull << K) / kBlockDim) * 6 * 16 + 2 * 2 * 6 * 16;
    static GridCache grid_cache;
    int grid_cap = 0;
    hipError_t e = persistent_grid(block_mfma64_kernel<K>, 256, lds, grid_cache, &grid_cap);
    if (e != hipSuccess) return e;
    const uint64_t ntiles = 1ull << (P.n - P.k);
    const unsigned grid = static_cast<unsigned>(ntiles < static_cast<uint64_t>(grid_cap) ? ntiles : grid_cap);
    hipLaunchKernelGGL(block_mfma64_kernel<K>, dim3(grid), dim3(256), lds, st, re, im, P, B, mat, tab, norm2);
    return hipGetLastError();
}

// Warp-specialised complex64 BLOCK kernel for tiles of 2^12 amplitudes: one
// workgroup per CU, 8 waves.  Waves 0-3 multiply (two 16-column groups each,
// one wave per SIMD; measured faster than two multiplying waves per SIMD),
// keeping only the upper half A = [Ur, -Ui] of the real form in registers
// (re' = A [re; im], im' = A [im; -re]); waves 4-5 load tiles straight into
// LDS with 16-byte global_load_lds (no VGPRs held for data in flight, two
// tiles ahead); waves 6-7 write finished tiles back and prepare the tile
// factors of the tile two ahead.  Four LDS tile
// buffers rotate: in iteration it, buffer it % 4 is multiplied, (it + 1) % 4
// is landing, (it + 2) % 4 is being filled and (it - 1) % 4 is written back.
// The LDS layout is float4-granular (a DMA lane writes 16 contiguous bytes):
// float4 slot P holds local float4 J = P ^ ((P >> 3) & 7), i.e. element j
// sits at swz4(j) = j ^ (((j >> 5) & 7) << 2).  Barriers are raw s_barrier
// with compiler memory clobbers; each role waits for its own counters first
// (compute and store: LDS; load: all but the newest tile's 16 DMA loads).
constexpr int kWsBuffers = 4;
constexpr int kWsOuterMax = kMaxQubits - 12;
constexpr size_t ws_lds_bytes() {
    return static_cast<size_t>(kWsBuffers) * 2 * 4096 * sizeof(float) + 2ull * 128 * 6 * 16 +
           static_cast<size_t>(kWsBuffers) * 2 * 6 * 16 + 2ull * kWsOuterMax * 6 * sizeof(double);
}

__device__ __forceinline__ uint32_t swz4(uint32_t j) { return j ^ (((j >> 5) & 7u) << 2); }

__device__ __forceinline__ void ws_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

__global__ __launch_bounds__(512) void block_ws_kernel(float* __restrict__ re, float* __restrict__ im,
                                                       const TilePass P, const BlockPass B,
                                                       const float* __restrict__ mat,
                                                       const double* __restrict__ tab,
                                                       const double* __restrict__ norm2) {
    constexpr int K = 12;
    constexpr uint32_t TS = 1u << K;
    constexpr int NB = kWsBuffers;
    constexpr int NLD = 128;                          // loader / storer lanes
    constexpr int NCHUNK = static_cast<int>(TS / 4 / NLD);   // DMA instructions per array and tile
    constexpr int NCOL = static_cast<int>(TS / kBlockDim);
    constexpr int NGROUP = NCOL / 16;
    constexpr int NZ = kBlockQubits + 1;
    constexpr int NCW = 4;                            // compute waves
    constexpr int NT = 512;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* s_buf = reinterpret_cast<float*>(lds_raw);                         // [NB][re | im][TS]
    double2* s_cf = reinterpret_cast<double2*>(s_buf + NB * 2 * TS);          // [2][NCOL][NZ]
    double2* s_tf = s_cf + 2 * NCOL * NZ;                                     // [NB][2][NZ]
    double* s_oa = reinterpret_cast<double*>(s_tf + NB * 2 * NZ);            // [2][nouter][NZ]
    const uint32_t tid = threadIdx.x;
    const int lane = static_cast<int>(tid & 63u);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
    const bool pre_any = B.pre_tab != 0, post_any = B.post_tab != 0;
    const double scale = B.renorm ? 1.0 / sqrt(norm2[0]) : 1.0;
    const uint64_t ntiles = 1ull << (P.n - P.k);
    const uint32_t niter = static_cast<uint64_t>(blockIdx.x) < ntiles
                               ? static_cast<uint32_t>((ntiles - 1 - blockIdx.x) / gridDim.x + 1) : 0u;
    auto tile_of = [&](uint32_t i) { return static_cast<uint64_t>(blockIdx.x) + static_cast<uint64_t>(i) * gridDim.x; };

    if (pre_any || post_any) {
        const double2* cf = reinterpret_cast<const double2*>(tab + B.tab_off);
        for (uint32_t i = tid; i < 2u * NCOL * NZ; i += NT) s_cf[i] = cf[i];
        const double* oa = tab + B.tab_off + 2 * NCOL * NZ * 2;
        for (uint32_t i = tid; i < 2u * P.nouter * NZ; i += NT) s_oa[i] = oa[i];
    }
    // tile factors exp(i sum of the tile's outer-bit angles), pre scaled (from LDS)
    auto tile_factors = [&](uint32_t i) {
        const uint32_t r = tid & 127u;
        if (r < 2 * NZ) {
            const int w = static_cast<int>(r) / NZ, j = static_cast<int>(r) % NZ;
            const uint64_t t = tile_of(i);
            const double* a = s_oa + static_cast<size_t>(w) * P.nouter * NZ;
            double x = 0.0;
            for (int q = 0; q < P.nouter; ++q)
                if ((t >> q) & 1ull) x += a[q * NZ + j];
            Z64 z = zexp(x);
            if (w == 0 && j == 0) { z.r *= scale; z.i *= scale; }
            s_tf[(i % NB) * 2 * NZ + r] = make_double2(z.r, z.i);
        }
    };
    // loader / storer lanes: float4 slot c * 128 + ll holds local float4
    // c * 128 + jll (the XOR only touches bits 0-2 of ll)
    const uint32_t ll = tid & 127u;
    const uint32_t jll = ll ^ ((ll >> 3) & 7u);
    const uint64_t goff = local_to_global<K>(4u * jll, P);
    auto issue_loads = [&](uint32_t i) {
        const uint64_t base = tile_base<K>(tile_of(i), P);
        float* br = s_buf + (i % NB) * 2 * TS;
#pragma unroll
        for (int c = 0; c < NCHUNK; ++c) {
            const uint64_t g = base + goff + local_to_global<K>(512u * static_cast<uint32_t>(c), P);
            const uint32_t slot0 = static_cast<uint32_t>(c * NLD + (wave - NCW) * 64);
            __builtin_amdgcn_global_load_lds(re + g, br + slot0 * 4, 16, 0, 0);
            __builtin_amdgcn_global_load_lds(im + g, br + TS + slot0 * 4, 16, 0, 0);
        }
    };
    constexpr int kVmcnt0 = 0x0F70, kVmcnt16 = 0x4F70, kLgkm0 = 0xC07F;

    // prologue: tiles 0 and 1 in flight, their factors ready, tile 0 landed
    if (wave == NCW || wave == NCW + 1) {
        if (niter > 0) issue_loads(0);
        if (niter > 1) { issue_loads(1); __builtin_amdgcn_s_waitcnt(kVmcnt16); }
        else __builtin_amdgcn_s_waitcnt(kVmcnt0);
    }
    __syncthreads();
    if (wave >= NCW + 2 && (pre_any || post_any)) {
        if (niter > 0) tile_factors(0);
        if (niter > 1) tile_factors(1);
    }
    __syncthreads();

    // compute-wave constants
    float afrag[2][16];   // rows 0-31 of the real form: [Ur, -Ui]
    uint32_t wmask = 0, rl_in = 0, rl_out = 0;
    if (wave < NCW) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int s2 = 0; s2 < 16; ++s2) afrag[r][s2] = mat[(16 * r + (lane & 15)) * 64 + 4 * s2 + (lane >> 4)];
    }
    uint32_t qmask = 0;
#pragma unroll
    for (int j = 0; j < kBlockQubits; ++j) qmask |= 1u << B.qloc[j];
    wmask = (TS - 1) & ~qmask;
    auto eloc = [&](int e) {
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < kBlockQubits; ++j) x |= static_cast<uint32_t>((e >> j) & 1) << B.qloc[j];
        return x;
    };
    auto col_local = [&](uint32_t col) {
        uint32_t colloc = 0, x = col, m = wmask;
        while (m) {
            const uint32_t low = m & (~m + 1);
            if (x & 1) colloc |= low;
            x >>= 1;
            m ^= low;
        }
        return colloc;
    };
    const int g = lane >> 4;
    {
        const uint32_t dl = col_local(static_cast<uint32_t>(lane & 15));
        rl_in = swz4(dl ^ eloc(g));
        rl_out = swz4(dl ^ eloc(4 * g));
    }

    for (uint32_t it = 0; it <= niter; ++it) {
        if (wave < NCW) {
            if (it < niter) {
                float* s_re = s_buf + (it % NB) * 2 * TS;
                float* s_im = s_re + TS;
                const double2* tf = s_tf + (it % NB) * 2 * NZ;
                for (int cg = wave; cg < NGROUP; cg += NCW) {
                    const uint32_t col = static_cast<uint32_t>(cg * 16 + (lane & 15));
                    const uint32_t ucg = swz4(col_local(static_cast<uint32_t>(cg * 16)));
                    float bfrag[16];
                    if (pre_any) {
                        float fr[8], fi[8];
                        Z64 z[NZ];
                        column_factors(s_cf + col * NZ, tf, z);
                        Z64 b0 = z[0];
                        if (g & 1) b0 = zmul(b0, z[1]);
                        if (g & 2) b0 = zmul(b0, z[2]);
                        element_factors(b0, z[3], z[4], z[5], fr, fi);
#pragma unroll
                        for (int m = 0; m < 8; ++m) {
                            const uint32_t j = rl_in ^ ucg ^ swz4(eloc(4 * m));
                            const float xr = s_re[j], xi = s_im[j];
                            bfrag[m] = xr * fr[m] - xi * fi[m];
                            bfrag[m + 8] = xr * fi[m] + xi * fr[m];
                        }
                    } else {
                        const float sc = static_cast<float>(scale);
#pragma unroll
                        for (int m = 0; m < 8; ++m) {
                            const uint32_t j = rl_in ^ ucg ^ swz4(eloc(4 * m));
                            bfrag[m] = s_re[j] * sc;
                            bfrag[m + 8] = s_im[j] * sc;
                        }
                    }
                    // acc[h]: re of rows 16 h + ..; acc[2 + h]: im (B' = [im; -re])
                    f4_t acc[4];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        acc[h] = f4_t{0.f, 0.f, 0.f, 0.f};
                        acc[2 + h] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int s2 = 0; s2 < 16; ++s2)
                            acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag[h][s2], bfrag[s2], acc[h], 0, 0, 0);
#pragma unroll
                        for (int s2 = 0; s2 < 8; ++s2) {
                            acc[2 + h] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag[h][s2], bfrag[8 + s2], acc[2 + h], 0, 0, 0);
                            acc[2 + h] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag[h][8 + s2], -bfrag[s2], acc[2 + h], 0, 0, 0);
                        }
                    }
                    if (post_any) {
                        float pr[8], pi2[8];
                        Z64 z[NZ];
                        column_factors(s_cf + NCOL * NZ + col * NZ, tf + NZ, z);
                        Z64 b0 = z[0];
                        if (g & 1) b0 = zmul(b0, z[3]);
                        if (g & 2) b0 = zmul(b0, z[4]);
                        element_factors(b0, z[1], z[2], z[5], pr, pi2);
#pragma unroll
                        for (int h = 0; h < 2; ++h)
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const float yr = acc[h][v], yi = acc[2 + h][v];
                                const int q = 4 * h + v;
                                const uint32_t j = rl_out ^ ucg ^ swz4(eloc(16 * h + v));
                                s_re[j] = yr * pr[q] - yi * pi2[q];
                                s_im[j] = yr * pi2[q] + yi * pr[q];
                            }
                    } else {
#pragma unroll
                        for (int h = 0; h < 2; ++h)
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const uint32_t j = rl_out ^ ucg ^ swz4(eloc(16 * h + v));
                                s_re[j] = acc[h][v];
                                s_im[j] = acc[2 + h][v];
                            }
                    }
                }
            }
            __builtin_amdgcn_s_waitcnt(kLgkm0);
        } else if (wave < NCW + 2) {
            if (it + 2 < niter) {
                issue_loads(it + 2);
                __builtin_amdgcn_s_waitcnt(kVmcnt16);
            } else {
                __builtin_amdgcn_s_waitcnt(kVmcnt0);
            }
        } else {
            // factors first: an LDS write after the stores would wait for them
            if ((pre_any || post_any) && it + 2 < niter) tile_factors(it + 2);
            if (it >= 1) {
                const uint32_t i = it - 1;
                const uint64_t base = tile_base<K>(tile_of(i), P);
                const float* br = s_buf + (i % NB) * 2 * TS;
#pragma unroll
                for (int c = 0; c < NCHUNK; ++c) {
                    const uint32_t slot = static_cast<uint32_t>(c * NLD) + ll;
                    const float4 x = *reinterpret_cast<const float4*>(br + 4 * slot);
                    const float4 y = *reinterpret_cast<const float4*>(br + TS + 4 * slot);
                    const uint64_t gi = base + goff + local_to_global<K>(512u * static_cast<uint32_t>(c), P);
                    *reinterpret_cast<float4*>(re + gi) = x;
                    *reinterpret_cast<float4*>(im + gi) = y;
                }
            }
            __builtin_amdgcn_s_waitcnt(kLgkm0);
        }
        ws_barrier();
    }
}

hipError_t launch_block_ws(float* re, float* im, const TilePass& P, const BlockPass& B, const float* mat,
                           const double* tab, const double* norm2, hipStream_t st) {
    const size_t lds = ws_lds_bytes();
    static GridCache grid_cache;
    int grid_cap = 0;
    hipError_t e = persistent_grid(block_ws_kernel, 512, lds, grid_cache, &grid_cap);
    if (e != hipSuccess) return e;
    const uint64_t ntiles = 1ull << (P.n - P.k);
    const unsigned grid = static_cast<unsigned>(ntiles < static_cast<uint64_t>(grid_cap) ? ntiles : grid_cap);
    hipLaunchKernelGGL(block_ws_kernel, dim3(grid), dim3(512), lds, st, re, im, P, B, mat, tab, norm2);
    return hipGetLastError();
}

// complex128 counterpart of block_ws_kernel for tiles of 2^11 amplitudes: the
// same roles and buffer rotation; a 16-byte DMA lane carries two doubles, so
// 16-byte slot P holds local pair J = P ^ ((P >> 3) & 7) and element j sits at
// swz2(j) = j ^ (((j >> 4) & 7) << 1).  The compute waves (one 16-column
// group each) run the FP64 MFMA code of block_mfma64_kernel.
constexpr int kWs64OuterMax = kMaxQubits - 11;
constexpr size_t ws64_lds_bytes() {
    return static_cast<size_t>(kWsBuffers) * 2 * 2048 * sizeof(double) + 2ull * 64 * 6 * 16 +
           static_cast<size_t>(kWsBuffers) * 2 * 6 * 16 + 2ull * kWs64OuterMax * 6 * sizeof(double);
}

__device__ __forceinline__ uint32_t swz2(uint32_t j) { return j ^ (((j >> 4) & 7u) << 1); }

__global__ __launch_bounds__(512) void block_ws64_kernel(double* __restrict__ re, double* __restrict__ im,
                                                         const TilePass P, const BlockPass B,
                                                         const double* __restrict__ mat,
                                                         const double* __restrict__ tab,
                                                         const double* __restrict__ norm2) {
    constexpr int K = 11;
    constexpr uint32_t TS = 1u << K;
    constexpr int NB = kWsBuffers;
    constexpr int NLD = 128;
    constexpr int NCHUNK = static_cast<int>(TS / 2 / NLD);
    constexpr int NCOL = static_cast<int>(TS / kBlockDim);
    constexpr int NGROUP = NCOL / 16;
    constexpr int NZ = kBlockQubits + 1;
    constexpr int NCW = 4;
    constexpr int NT = 512;
    static_assert(NGROUP == NCW, "one group per compute wave");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double* s_buf = reinterpret_cast<double*>(lds_raw);                       // [NB][re | im][TS]
    double2* s_cf = reinterpret_cast<double2*>(s_buf + NB * 2 * TS);          // [2][NCOL][NZ]
    double2* s_tf = s_cf + 2 * NCOL * NZ;                                     // [NB][2][NZ]
    double* s_oa = reinterpret_cast<double*>(s_tf + NB * 2 * NZ);            // [2][nouter][NZ]
    const uint32_t tid = threadIdx.x;
    const int lane = static_cast<int>(tid & 63u);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
    const bool pre_any = B.pre_tab != 0, post_any = B.post_tab != 0;
    const double scale = B.renorm ? 1.0 / sqrt(norm2[0]) : 1.0;
    const uint64_t ntiles = 1ull << (P.n - P.k);
    const uint32_t niter = static_cast<uint64_t>(blockIdx.x) < ntiles
                               ? static_cast<uint32_t>((ntiles - 1 - blockIdx.x) / gridDim.x + 1) : 0u;
    auto tile_of = [&](uint32_t i) { return static_cast<uint64_t>(blockIdx.x) + static_cast<uint64_t>(i) * gridDim.x; };

    if (pre_any || post_any) {
        const double2* cf = reinterpret_cast<const double2*>(tab + B.tab_off);
        for (uint32_t i = tid; i < 2u * NCOL * NZ; i += NT) s_cf[i] = cf[i];
        const double* oa = tab + B.tab_off + 2 * NCOL * NZ * 2;
        for (uint32_t i = tid; i < 2u * P.nouter * NZ; i += NT) s_oa[i] = oa[i];
    }
    auto tile_factors = [&](uint32_t i) {
        const uint32_t r = tid & 127u;
        if (r < 2 * NZ) {
            const int w = static_cast<int>(r) / NZ, j = static_cast<int>(r) % NZ;
            const uint64_t t = tile_of(i);
            const double* a = s_oa + static_cast<size_t>(w) * P.nouter * NZ;
            double x = 0.0;
            for (int q = 0; q < P.nouter; ++q)
                if ((t >> q) & 1ull) x += a[q * NZ + j];
            Z64 z = zexp(x);
            if (w == 0 && j == 0) { z.r *= scale; z.i *= scale; }
            s_tf[(i % NB) * 2 * NZ + r] = make_double2(z.r, z.i);
        }
    };
    const uint32_t ll = tid & 127u;
    const uint32_t jll = ll ^ ((ll >> 3) & 7u);
    const uint64_t goff = local_to_global<K>(2u * jll, P);
    auto issue_loads = [&](uint32_t i) {
        const uint64_t base = tile_base<K>(tile_of(i), P);
        double* br = s_buf + (i % NB) * 2 * TS;
#pragma unroll
        for (int c = 0; c < NCHUNK; ++c) {
            const uint64_t g = base + goff + local_to_global<K>(256u * static_cast<uint32_t>(c), P);
            const uint32_t slot0 = static_cast<uint32_t>(c * NLD + (wave - NCW) * 64);
            __builtin_amdgcn_global_load_lds(re + g, br + slot0 * 2, 16, 0, 0);
            __builtin_amdgcn_global_load_lds(im + g, br + TS + slot0 * 2, 16, 0, 0);
        }
    };
    constexpr int kVmcnt0 = 0x0F70, kVmcnt16 = 0x4F70, kLgkm0 = 0xC07F;

    if (wave == NCW || wave == NCW + 1) {
        if (niter > 0) issue_loads(0);
        if (niter > 1) { issue_loads(1); __builtin_amdgcn_s_waitcnt(kVmcnt16); }
        else __builtin_amdgcn_s_waitcnt(kVmcnt0);
    }
    __syncthreads();
    if (wave >= NCW + 2 && (pre_any || post_any)) {
        if (niter > 0) tile_factors(0);
        if (niter > 1) tile_factors(1);
    }
    __syncthreads();

    double afrag[2][16];
    if (wave < NCW) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int s2 = 0; s2 < 16; ++s2) afrag[r][s2] = mat[(16 * r + (lane & 15)) * 64 + 4 * s2 + (lane >> 4)];
    }
    uint32_t qmask = 0;
#pragma unroll
    for (int j = 0; j < kBlockQubits; ++j) qmask |= 1u << B.qloc[j];
    const uint32_t wmask = (TS - 1) & ~qmask;
    auto eloc = [&](int e) {
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < kBlockQubits; ++j) x |= static_cast<uint32_t>((e >> j) & 1) << B.qloc[j];
        return x;
    };
    auto col_local = [&](uint32_t col) {
        uint32_t colloc = 0, x = col, m = wmask;
        while (m) {
            const uint32_t low = m & (~m + 1);
            if (x & 1) colloc |= low;
            x >>= 1;
            m ^= low;
        }
        return colloc;
    };
    const int g = lane >> 4;
    const uint32_t dl = col_local(static_cast<uint32_t>(lane & 15));
    const uint32_t rl = swz2(dl ^ eloc(g));   // rows e = g + 4 m (in) and e = g + 4 v + 16 h (out)

    for (uint32_t it = 0; it <= niter; ++it) {
        if (wave < NCW) {
            if (it < niter) {
                double* s_re = s_buf + (it % NB) * 2 * TS;
                double* s_im = s_re + TS;
                const double2* tf = s_tf + (it % NB) * 2 * NZ;
                const int cg = wave;
                const uint32_t col = static_cast<uint32_t>(cg * 16 + (lane & 15));
                const uint32_t ucg = swz2(col_local(static_cast<uint32_t>(cg * 16)));
                double br[8], bi[8];
                if (pre_any) {
                    Z64 z[NZ];
                    column_factors(s_cf + col * NZ, tf, z);
                    Z64 b0 = z[0];
                    if (g & 1) b0 = zmul(b0, z[1]);
                    if (g & 2) b0 = zmul(b0, z[2]);
                    Z64 f[8];
                    f[0] = b0;
                    f[1] = zmul(f[0], z[3]);
                    f[2] = zmul(f[0], z[4]);
                    f[3] = zmul(f[1], z[4]);
#pragma unroll
                    for (int m = 0; m < 4; ++m) f[4 + m] = zmul(f[m], z[5]);
#pragma unroll
                    for (int m = 0; m < 8; ++m) {
                        const uint32_t j = rl ^ ucg ^ swz2(eloc(4 * m));
                        const double xr = s_re[j], xi = s_im[j];
                        br[m] = xr * f[m].r - xi * f[m].i;
                        bi[m] = xr * f[m].i + xi * f[m].r;
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < 8; ++m) {
                        const uint32_t j = rl ^ ucg ^ swz2(eloc(4 * m));
                        br[m] = s_re[j] * scale;
                        bi[m] = s_im[j] * scale;
                    }
                }
                d4_t acc_r[2], acc_i[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    acc_r[r] = d4_t{0.0, 0.0, 0.0, 0.0};
                    acc_i[r] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int s2 = 0; s2 < 8; ++s2) {
                        acc_r[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][s2], br[s2], acc_r[r], 0, 0, 0);
                        acc_i[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][s2], bi[s2], acc_i[r], 0, 0, 0);
                    }
#pragma unroll
                    for (int s2 = 0; s2 < 8; ++s2) {
                        acc_r[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][8 + s2], bi[s2], acc_r[r], 0, 0, 0);
                        acc_i[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[r][8 + s2], -br[s2], acc_i[r], 0, 0, 0);
                    }
                }
                if (post_any) {
                    Z64 z[NZ];
                    column_factors(s_cf + NCOL * NZ + col * NZ, tf + NZ, z);
                    Z64 b0 = z[0];
                    if (g & 1) b0 = zmul(b0, z[1]);
                    if (g & 2) b0 = zmul(b0, z[2]);
                    Z64 f[8];
                    f[0] = b0;
                    f[1] = zmul(f[0], z[3]);
                    f[2] = zmul(f[0], z[4]);
                    f[3] = zmul(f[1], z[4]);
#pragma unroll
                    for (int v = 0; v < 4; ++v) f[4 + v] = zmul(f[v], z[5]);
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const double yr = acc_r[h][v], yi = acc_i[h][v];
                            const Z64 fq = f[v + 4 * h];
                            const uint32_t j = rl ^ ucg ^ swz2(eloc(16 * h + 4 * v));
                            s_re[j] = yr * fq.r - yi * fq.i;
                            s_im[j] = yr * fq.i + yi * fq.r;
                        }
                } else {
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const uint32_t j = rl ^ ucg ^ swz2(eloc(16 * h + 4 * v));
                            s_re[j] = acc_r[h][v];
                            s_im[j] = acc_i[h][v];
                        }
                }
            }
            __builtin_amdgcn_s_waitcnt(kLgkm0);
        } else if (wave < NCW + 2) {
            if (it + 2 < niter) {
                issue_loads(it + 2);
                __builtin_amdgcn_s_waitcnt(kVmcnt16);
            } else {
                __builtin_amdgcn_s_waitcnt(kVmcnt0);
            }
        } else {
            if ((pre_any || post_any) && it + 2 < niter) tile_factors(it + 2);
            if (it >= 1) {
                const uint32_t i = it - 1;
                const uint64_t base = tile_base<K>(tile_of(i), P);
                const double* br = s_buf + (i % NB) * 2 * TS;
#pragma unroll
                for (int c = 0; c < NCHUNK; ++c) {
                    const uint32_t slot = static_cast<uint32_t>(c * NLD) + ll;
                    const double2 x = *reinterpret_cast<const double2*>(br + 2 * slot);
                    const double2 y = *reinterpret_cast<const double2*>(br + TS + 2 * slot);
                    const uint64_t gi = base + goff + local_to_global<K>(256u * static_cast<uint32_t>(c), P);
                    *reinterpret_cast<double2*>(re + gi) = x;
                    *reinterpret_cast<double2*>(im + gi) = y;
                }
            }
            __builtin_amdgcn_s_waitcnt(kLgkm0);
        }
        ws_barrier();
    }
}

hipError_t launch_block_ws64(double* re, double* im, const TilePass& P, const BlockPass& B, const double* mat,
                             const double* tab, const double* norm2, hipStream_t st) {
    const size_t lds = ws64_lds_bytes();
    static GridCache grid_cache;
    int grid_cap = 0;
    hipError_t e = persistent_grid(block_ws64_kernel, 512, lds, grid_cache, &grid_cap);
    if (e != hipSuccess) return e;
    const uint64_t ntiles = 1ull << (P.n - P.k);
    const unsigned grid = static_cast<unsigned>(ntiles < static_cast<uint64_t>(grid_cap) ? ntiles : grid_cap);
    hipLaunchKernelGGL(block_ws64_kernel, dim3(grid), dim3(512), lds, st, re, im, P, B, mat, tab, norm2);
    return hipGetLastError();
}

template <int K>
hipError_t launch_block_k(float* re, float* im, const TilePass& P, const BlockPass& B, const float* mat,
                          const double* tab, const double* norm2, hipStream_t st) {
    // the tile (re, im), the pre / post column factors (6 complex FP64 per
    // column) and two buffers of the tile's 2 x 6 factors
    const size_t lds = 2ull * (1ull << K) * sizeof(float) + 2ull * ((1ull << K) / kBlockDim) * 6 * 16 + 2 * 2 * 6 * 16;
    static GridCache grid_cache;
    int grid_cap = 0;
    hipError_t e = persistent_grid(block_mfma_kernel<K>, 256, lds, grid_cache, &grid_cap);
    if (e != hipSuccess) return e;
    const uint64_t ntiles = 1ull << (P.n - P.k);
    const unsigned grid = static_cast<unsigned>(ntiles < static_cast<uint64_t>(grid_cap) ? ntiles : grid_cap);
    hipLaunchKernelGGL(block_mfma_kernel<K>, dim3(grid), dim3(256), lds, st, re, im, P, B, mat, tab, norm2);
    return hipGetLastError();
}

// ---------------------------------------------------------------- (c) permutations
__device__ __forceinline__ uint32_t rev_bits32(uint32_t x, int nbits) {
    return nbits == 0 ? 0u : (__brev(x) >> (32 - nbits));
}

__device__ __forceinline__ uint64_t rev_bits64(uint64_t x, int nbits) {
    return nbits == 0 ? 0ull : (__brevll(x) >> (64 - nbits));
}

// Index i = (h: top B bits | m: middle M bits | l: low B bits); its reversal is
// (rev l | rev m | rev h).  Tile m (with m <= rev m) is exchanged with tile
// rev m through LDS, re and im together: complex-interleaved W x W tiles
// (padded rows), 16-byte vectors along l, one load phase and one store phase
// per tile pair.  Pure data movement: bit-exact.  The block order keeps the
// pages in use few.  Tile m's rows lie
// 2^(n-B) amplitudes apart and each page holds a run of consecutive m; the
// partner tile rev m lies in the page run of m's low bits.  Consecutive
// iterations u therefore step through m's middle bits first (u bits [0, c)
// -> m bits [a, a + c), u bits [c, c + a) -> m bits [0, a), the rest above),
// so both the blocks in flight and their partners share a few pages.  The
// grid is persistent (blocks loop over u).
template <typename T, int B>
__global__ __launch_bounds__(256) void bitrev_joint_kernel(T* __restrict__ re, T* __restrict__ im, int n) {
    constexpr int W = 1 << B;
    constexpr int VN = 16 / static_cast<int>(sizeof(T));
    constexpr int RV = W / VN;
    using V = typename Vec<T>::type;
    struct C2 {
        T r, i;
    };
    __shared__ C2 ta[W][W + 1], tb[W][W + 1];
    const int M = n - 2 * B;
    const int hs = n - B;
    const int c = M < 8 ? M : 8;
    const int a = (M - c) / 2;
    const uint64_t cm = (1ull << c) - 1, am = (1ull << a) - 1;
    for (uint64_t u = blockIdx.x; u < (1ull << M); u += gridDim.x) {
        const uint64_t m = ((u & cm) << a) | ((u >> c) & am) | ((u >> (c + a)) << (a + c));
        const uint64_t rm = rev_bits64(m, M);
        if (m > rm) continue;   // uniform across the block
        const bool self = (m == rm);
        for (int e = threadIdx.x; e < W * RV; e += 256) {
            const int h = e / RV, l0 = (e % RV) * VN;
            const uint64_t ia = (static_cast<uint64_t>(h) << hs) | (m << B) | static_cast<uint64_t>(l0);
            const V ra = *reinterpret_cast<const V*>(re + ia);
            const V ia_ = *reinterpret_cast<const V*>(im + ia);
            V rb, ib_;
            if (!self) {
                const uint64_t ib = (static_cast<uint64_t>(h) << hs) | (rm << B) | static_cast<uint64_t>(l0);
                rb = *reinterpret_cast<const V*>(re + ib);
                ib_ = *reinterpret_cast<const V*>(im + ib);
            }
            const T* xr = reinterpret_cast<const T*>(&ra);
            const T* xi = reinterpret_cast<const T*>(&ia_);
#pragma unroll
            for (int j = 0; j < VN; ++j) ta[h][l0 + j] = C2{xr[j], xi[j]};
            if (!self) {
                const T* yr = reinterpret_cast<const T*>(&rb);
                const T* yi = reinterpret_cast<const T*>(&ib_);
#pragma unroll
                for (int j = 0; j < VN; ++j) tb[h][l0 + j] = C2{yr[j], yi[j]};
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < W * RV; e += 256) {
            const int h = e / RV, l0 = (e % RV) * VN;
            const int rh = static_cast<int>(rev_bits32(static_cast<uint32_t>(h), B));
            V oar, oai, obr, obi;
            T* par = reinterpret_cast<T*>(&oar);
            T* pai = reinterpret_cast<T*>(&oai);
            T* pbr = reinterpret_cast<T*>(&obr);
            T* pbi = reinterpret_cast<T*>(&obi);
#pragma unroll
            for (int j = 0; j < VN; ++j) {
                const int rl = static_cast<int>(rev_bits32(static_cast<uint32_t>(l0 + j), B));
                const C2 x = self ? ta[rl][rh] : tb[rl][rh];
                const C2 y = ta[rl][rh];
                par[j] = x.r;
                pai[j] = x.i;
                pbr[j] = y.r;
                pbi[j] = y.i;
            }
            const uint64_t ia = (static_cast<uint64_t>(h) << hs) | (m << B) | static_cast<uint64_t>(l0);
            *reinterpret_cast<V*>(re + ia) = oar;
            *reinterpret_cast<V*>(im + ia) = oai;
            if (!self) {
                const uint64_t ib = (static_cast<uint64_t>(h) << hs) | (rm << B) | static_cast<uint64_t>(l0);
                *reinterpret_cast<V*>(re + ib) = obr;
                *reinterpret_cast<V*>(im + ib) = obi;
            }
        }
        __syncthreads();   // the next tile pair reuses the buffers
    }
}

// Reversal of the bits in `mask` (partial_bit_reverse_state, state.mojo:252-301)
// or of all bits: exchange i and perm(i) when i < perm(i).
template <typename T>
__global__ __launch_bounds__(256) void bitrev_simple_kernel(T* __restrict__ re, T* __restrict__ im, uint64_t size,
                                                            uint64_t mask, int n) {
    int pos[64];
    int k = 0;
    for (int b = 0; b < n; ++b)
        if ((mask >> b) & 1) pos[k++] = b;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < size; i += stride) {
        uint64_t j = i & ~mask;
        for (int b = 0; b < k; ++b) j |= ((i >> pos[b]) & 1ull) << pos[k - 1 - b];
        if (i < j) {
            const T tr = re[i], ti = im[i];
            re[i] = re[j];
            im[i] = im[j];
            re[j] = tr;
            im[j] = ti;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void swap_kernel(T* __restrict__ re, T* __restrict__ im, uint64_t count, int lo,
                                                   int hi, int a, int b) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; w < count; w += stride) {
        const uint64_t base = insert_zero64(insert_zero64(w, lo), hi);
        const uint64_t i = base | (1ull << a), j = base | (1ull << b);
        const T tr = re[i], ti = im[i];
        re[i] = re[j];
        im[i] = im[j];
        re[j] = tr;
        im[j] = ti;
    }
}

// Dense 2^m block (apply_unitary / apply_controlled_unitary,
// transformations_unitary.mojo:22-139): a block handles G = 256 / dim groups;
// each group's 2^m inputs are staged in LDS, each thread forms output rows.
struct UnitaryArgs {
    uint64_t ngroups;
    uint64_t controls;
    int32_t target, m, npos;
    int32_t pos[kMaxQubits + 1];  // ascending positions of block and control bits
};

__device__ __forceinline__ uint64_t unitary_base(uint64_t grp, const UnitaryArgs& A) {
    uint64_t base = grp;
    for (int p = 0; p < A.npos; ++p) base = insert_zero64(base, A.pos[p]);
    return base | A.controls;
}

template <typename T>
__global__ __launch_bounds__(256) void unitary_kernel(T* __restrict__ re, T* __restrict__ im, const UnitaryArgs A,
                                                      const double* __restrict__ u) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int dim = 1 << A.m;
    const int G = dim >= 256 ? 1 : 256 / dim;
    T* x_re = reinterpret_cast<T*>(lds_raw);
    T* x_im = x_re + G * dim;
    for (uint64_t g0 = static_cast<uint64_t>(blockIdx.x) * G; g0 < A.ngroups;
         g0 += static_cast<uint64_t>(gridDim.x) * G) {
        for (int e = threadIdx.x; e < G * dim; e += 256) {
            const int gl = e / dim, k = e % dim;
            const uint64_t grp = g0 + gl;
            if (grp >= A.ngroups) continue;
            const uint64_t idx = unitary_base(grp, A) + (static_cast<uint64_t>(k) << A.target);
            x_re[e] = re[idx];
            x_im[e] = im[idx];
        }
        __syncthreads();
        T o_re[4], o_im[4];  // up to 4 rows per thread (dim <= 1024)
        int cnt = 0;
        for (int e = threadIdx.x; e < G * dim; e += 256, ++cnt) {
            const int gl = e / dim, row = e % dim;
            T sr = 0, si = 0;
            for (int col = 0; col < dim; ++col) {
                const T ur = static_cast<T>(u[2 * (row * dim + col)]);
                const T ui = static_cast<T>(u[2 * (row * dim + col) + 1]);
                const T vr = x_re[gl * dim + col], vi = x_im[gl * dim + col];
                sr += ur * vr - ui * vi;
                si += ur * vi + ui * vr;
            }
            o_re[cnt] = sr;
            o_im[cnt] = si;
        }
        __syncthreads();  // every input of the group is read before any output lands
        cnt = 0;
        for (int e = threadIdx.x; e < G * dim; e += 256, ++cnt) {
            const int gl = e / dim, row = e % dim;
            const uint64_t grp = g0 + gl;
            if (grp >= A.ngroups) continue;
            const uint64_t idx = unitary_base(grp, A) + (static_cast<uint64_t>(row) << A.target);
            re[idx] = o_re[cnt];
            im[idx] = o_im[cnt];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void reset_kernel(T* __restrict__ re, T* __restrict__ im, uint64_t size) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < size; i += stride) {
        re[i] = i == 0 ? T(1) : T(0);
        im[i] = T(0);
    }
}

// SplitMix64 output number `idx` (1-based) of the stream started at `seed`.
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + idx * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// re, im ~ N(0,1): uniforms u = top 53 bits of outputs 2i+1, 2i+2;
// Box-Muller with u1 = 1 - u (in (0, 1]).
template <typename T>
__global__ __launch_bounds__(256) void random_fill_kernel(T* __restrict__ re, T* __restrict__ im, uint64_t size,
                                                          uint64_t seed) {
    const double two_pi = 6.283185307179586;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < size; i += stride) {
        const double u1 = 1.0 - static_cast<double>(splitmix64_at(seed, 2 * i + 1) >> 11) * 0x1p-53;
        const double u2 = static_cast<double>(splitmix64_at(seed, 2 * i + 2) >> 11) * 0x1p-53;
        const double r = sqrt(-2.0 * log(u1));
        re[i] = static_cast<T>(r * cos(two_pi * u2));
        im[i] = static_cast<T>(r * sin(two_pi * u2));
    }
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void convert_kernel(const S* __restrict__ sre, const S* __restrict__ sim,
                                                      D* __restrict__ dre, D* __restrict__ dim, uint64_t size) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < size; i += stride) {
        dre[i] = static_cast<D>(sre[i]);
        dim[i] = static_cast<D>(sim[i]);
    }
}

// Scale to unit norm with the squared norm left on the device by a reduction.
template <typename T>
__global__ __launch_bounds__(256) void normalize_kernel(T* __restrict__ re, T* __restrict__ im, uint64_t size,
                                                        const double* __restrict__ norm2) {
    const T sc = static_cast<T>(1.0 / sqrt(norm2[0]));
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < size; i += stride) {
        re[i] *= sc;
        im[i] *= sc;
    }
}

// Streaming form of the library's in-place read + write (bfly_scale), the
// measured HBM ceiling of a pass: 16-byte vectors of re and im.
template <typename T>
struct SVec;
template <>
struct SVec<double> { typedef double type __attribute__((ext_vector_type(2))); static constexpr int N = 2; };
template <>
struct SVec<float> { typedef float type __attribute__((ext_vector_type(4))); static constexpr int N = 4; };
// One 16-byte vector of re and of im per thread and a block per 256 vectors
// (full grid): measured on MI355X for an in-place read + write of 2 x 8 GiB,
// plain accesses 5.75 TB/s; non-temporal 5.62 TB/s; 2 vectors per thread
// 5.33 (plain) / 5.61 (non-temporal); 8 per thread 5.21 / 5.45 TB/s.
constexpr int kStreamUnroll = 1;

template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(T* __restrict__ re, T* __restrict__ im, uint64_t size, double s) {
    using V = typename SVec<T>::type;
    constexpr int VN = SVec<T>::N;
    const T sc = static_cast<T>(s);
    const uint64_t nvec = size / VN;
    V* vr = reinterpret_cast<V*>(re);
    V* vi = reinterpret_cast<V*>(im);
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x * kStreamUnroll;
    for (uint64_t v0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x * kStreamUnroll + threadIdx.x; v0 < nvec;
         v0 += step) {
        V a[kStreamUnroll], b[kStreamUnroll];
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) {
            const uint64_t v = v0 + static_cast<uint64_t>(u) * blockDim.x;
            if (v < nvec) {
                a[u] = vr[v];
                b[u] = vi[v];
            }
        }
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) {
            const uint64_t v = v0 + static_cast<uint64_t>(u) * blockDim.x;
            if (v < nvec) {
                vr[v] = a[u] * sc;
                vi[v] = b[u] * sc;
            }
        }
    }
    // the last size % VN elements (states smaller than one vector)
    if (blockIdx.x == 0 && threadIdx.x < size - nvec * VN) {
        const uint64_t i = nvec * VN + threadIdx.x;
        re[i] *= sc;
        im[i] *= sc;
    }
}

// Scatter the bits of x into the positions of `mask` (ascending).
__device__ __forceinline__ uint64_t deposit(uint64_t x, uint64_t mask) {
    uint64_t r = 0;
    while (mask) {
        const uint64_t low = mask & (~mask + 1);
        if (x & 1) r |= low;
        x >>= 1;
        mask ^= low;
    }
    return r;
}

// Item of global work unit u: the last item whose `begin` is <= u.
__device__ __forceinline__ const OpItem& find_item(const OpItem& one, const OpItem* __restrict__ items, int nitems,
                                                   uint64_t u) {
    if (!items) return one;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].begin <= u) lo = mid;
        else hi = mid - 1;
    }
    return items[lo];
}

// PHASE_FLIP: negate the 2^(n - |mask|) amplitudes with (i & mask) == value
// (sign-bit flips, exact), for every item of the launch.
template <typename T>
__global__ __launch_bounds__(256) void flip_kernel(const OpItem one, const OpItem* __restrict__ items, int nitems,
                                                   uint64_t total) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t u = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < total; u += stride) {
        const OpItem& it = find_item(one, items, nitems, u);
        const uint64_t idx = deposit(u - it.begin, it.others) | it.value;
        T* re = static_cast<T*>(it.re);
        T* im = static_cast<T*>(it.im);
        re[idx] = -re[idx];
        im[idx] = -im[idx];
    }
}

// ---------------------------------------------------------------- (d) reductions
// TwoSum accumulation (Knuth): s + c carries the sum with its rounding error.
__device__ __forceinline__ void two_sum_add(double& s, double& c, double x) {
    const double t = s + x;
    const double bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
}

__device__ __forceinline__ void acc_merge(double& s, double& c, double s2, double c2) {
    two_sum_add(s, c, s2);
    c += c2;
}

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

template <typename T, int MODE>
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(const T* __restrict__ are, const T* __restrict__ aim,
                                                             const T* __restrict__ bre, const T* __restrict__ bim,
                                                             uint64_t size, uint64_t mask, uint64_t value,
                                                             double* __restrict__ partial) {
    constexpr int NC = MODE == R_INNER ? 2 : 1;
    __shared__ double sh[NC][2][kRedThreads];
    double s[NC], c[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) { s[q] = 0; c[q] = 0; }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRedThreads;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRedThreads + threadIdx.x; i < size; i += stride) {
        const double ar = are[i], ai = aim[i];
        if (MODE == R_NORM2) {
            two_sum_add(s[0], c[0], ar * ar + ai * ai);
        } else if (MODE == R_PROB_MASK) {
            if ((i & mask) == value) two_sum_add(s[0], c[0], ar * ar + ai * ai);
        } else if (MODE == R_DIFF2) {
            const double dr = ar - static_cast<double>(bre[i]), di = ai - static_cast<double>(bim[i]);
            two_sum_add(s[0], c[0], dr * dr + di * di);
        } else {
            const double br = bre[i], bi = bim[i];
            two_sum_add(s[0], c[0], ar * br + ai * bi);
            two_sum_add(s[NC - 1], c[NC - 1], ar * bi - ai * br);
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) { sh[q][0][threadIdx.x] = s[q]; sh[q][1][threadIdx.x] = c[q]; }
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(w)) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                double ss = sh[q][0][threadIdx.x], cc = sh[q][1][threadIdx.x];
                acc_merge(ss, cc, sh[q][0][threadIdx.x + w], sh[q][1][threadIdx.x + w]);
                sh[q][0][threadIdx.x] = ss;
                sh[q][1][threadIdx.x] = cc;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            partial[2 * (q * gridDim.x + blockIdx.x)] = sh[q][0][0];
            partial[2 * (q * gridDim.x + blockIdx.x) + 1] = sh[q][1][0];
        }
    }
}

// Final stage: channel q's nb partial (s, c) pairs -> out[q] = s + c.
__global__ __launch_bounds__(kRedThreads) void reduce_final_kernel(const double* __restrict__ partial, int nb, int nch,
                                                                   double* __restrict__ out) {
    __shared__ double sh[2][kRedThreads];
    for (int q = 0; q < nch; ++q) {
        double s = 0, c = 0;
        for (int i = threadIdx.x; i < nb; i += kRedThreads)
            acc_merge(s, c, partial[2 * (q * nb + i)], partial[2 * (q * nb + i) + 1]);
        sh[0][threadIdx.x] = s;
        sh[1][threadIdx.x] = c;
        __syncthreads();
        for (int w = kRedThreads / 2; w > 0; w >>= 1) {
            if (threadIdx.x < static_cast<unsigned>(w)) {
                double ss = sh[0][threadIdx.x], cc = sh[1][threadIdx.x];
                acc_merge(ss, cc, sh[0][threadIdx.x + w], sh[1][threadIdx.x + w]);
                sh[0][threadIdx.x] = ss;
                sh[1][threadIdx.x] = cc;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = sh[0][0] + sh[1][0];
        __syncthreads();
    }
}

// DIFFUSE over every qubit.  Chunk c of an item holds amplitudes
// [c 2^chunk_bits, (c + 1) 2^chunk_bits); an item has at most 1024 chunks.
// Summation order, shared by the partial pass and a chained apply pass:
// thread t takes the 16-byte vectors t, t + 256, t + 512, ... of the chunk
// and each vector's amplitudes in order (TwoSum per thread); lane 0 of each
// wave merges its wave's pairs by shuffles at distances 32, 16, ..., 1, and
// thread 0 merges the four waves in order.  Every workgroup of the apply
// pass merges the item's chunk pairs in the same fixed order (thread t takes
// chunks t, t + 256, ...; then the same workgroup merge), so all chunks use
// the same mean and the bits do not depend on the other items of a batch.
constexpr int kDiffChunkThreads = 256;
constexpr int kDiffWaves = kDiffChunkThreads / 64;
constexpr int kDiffVecMinBits = 6;   // smaller chunks go amplitude by amplitude

template <typename T>
struct DiffVec;
template <>
struct DiffVec<float> {
    using V = float4;
    static constexpr int E = 4;
};
template <>
struct DiffVec<double> {
    using V = double2;
    static constexpr int E = 2;
};

// v = (s_re, c_re, s_im, c_im); the merged result is valid in thread 0.
__device__ __forceinline__ void diff_block_merge(double (&v)[4], double (*sh)[4]) {
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) {
        double o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) o[q] = __shfl_down(v[q], w, 64);
        acc_merge(v[0], v[1], o[0], o[1]);
        acc_merge(v[2], v[3], o[2], o[3]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) sh[threadIdx.x >> 6][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 1; k < kDiffWaves; ++k) {
            acc_merge(v[0], v[1], sh[k][0], sh[k][1]);
            acc_merge(v[2], v[3], sh[k][2], sh[k][3]);
        }
    }
    __syncthreads();
}

// Chunk c's merged pair -> partial[4c .. 4c + 3].
__device__ __forceinline__ void diff_store_chunk(double (&v)[4], double* partial, uint64_t c, double (*sh)[4]) {
    diff_block_merge(v, sh);
    if (threadIdx.x == 0) {
        reinterpret_cast<double2*>(partial)[2 * c] = make_double2(v[0], v[1]);
        reinterpret_cast<double2*>(partial)[2 * c + 1] = make_double2(v[2], v[3]);
    }
}

// Pass 1: chunk sums (items whose sums the previous diffuser's apply pass
// already wrote are skipped).
template <typename T>
__global__ __launch_bounds__(kDiffChunkThreads) void diffuse_partial_kernel(const OpItem one,
                                                                           const OpItem* __restrict__ items,
                                                                           int nitems, uint64_t total) {
    using V = typename DiffVec<T>::V;
    constexpr int E = DiffVec<T>::E;
    __shared__ double sh[kDiffWaves][4];
    for (uint64_t u = blockIdx.x; u < total; u += gridDim.x) {
        const OpItem& it = find_item(one, items, nitems, u);
        if (it.flags & kChainReady) continue;
        const uint64_t c = u - it.begin;
        const uint64_t len = 1ull << it.chunk_bits;
        const T* re = static_cast<const T*>(it.re) + (c << it.chunk_bits);
        const T* im = static_cast<const T*>(it.im) + (c << it.chunk_bits);
        double v[4] = {0, 0, 0, 0};
        if (it.chunk_bits >= kDiffVecMinBits) {
            const V* rv = reinterpret_cast<const V*>(re);
            const V* iv = reinterpret_cast<const V*>(im);
#pragma unroll 4
            for (uint64_t k = threadIdx.x; k < len / E; k += kDiffChunkThreads) {
                const V xr = rv[k], xi = iv[k];
                const T* a = reinterpret_cast<const T*>(&xr);
                const T* b = reinterpret_cast<const T*>(&xi);
#pragma unroll
                for (int j = 0; j < E; ++j) {
                    two_sum_add(v[0], v[1], static_cast<double>(a[j]));
                    two_sum_add(v[2], v[3], static_cast<double>(b[j]));
                }
            }
        } else {
            for (uint64_t i = threadIdx.x; i < len; i += kDiffChunkThreads) {
                two_sum_add(v[0], v[1], static_cast<double>(re[i]));
                two_sum_add(v[2], v[3], static_cast<double>(im[i]));
            }
        }
        diff_store_chunk(v, it.partial, c, sh);
    }
}

// Pass 2: every chunk merges the item's pairs into 2 mean and updates its
// amplitudes: variant 0: a <- a - 2 mean; variant 1 (the reference's
// classical diffuser): re <- 2 mean_re - re, im <- -2 mean_im - im.  A
// chained pass then negates the amplitudes with (i & fmask) == fvalue and
// sums the results, in pass 1's order, into the next diffuser's partials.
template <typename T>
__device__ __forceinline__ void diff_update(T& x_re, T& x_im, double ar, double ai, int variant, bool flip,
                                            uint64_t idx, uint64_t fmask, uint64_t fvalue, bool next,
                                            double (&v)[4]) {
    const double xr = x_re, xi = x_im;
    T yr, yi;
    if (variant == 0) {
        yr = static_cast<T>(xr - ar);
        yi = static_cast<T>(xi - ai);
    } else {
        yr = static_cast<T>(ar - xr);
        yi = static_cast<T>(-ai - xi);
    }
    if (flip && (idx & fmask) == fvalue) {
        yr = -yr;
        yi = -yi;
    }
    x_re = yr;
    x_im = yi;
    if (next) {
        two_sum_add(v[0], v[1], static_cast<double>(yr));
        two_sum_add(v[2], v[3], static_cast<double>(yi));
    }
}

template <typename T>
__global__ __launch_bounds__(kDiffChunkThreads) void diffuse_apply_kernel(const OpItem one,
                                                                         const OpItem* __restrict__ items,
                                                                         int nitems, uint64_t total) {
    using V = typename DiffVec<T>::V;
    constexpr int E = DiffVec<T>::E;
    __shared__ double sh[kDiffWaves][4];
    __shared__ double mean[2];
    for (uint64_t u = blockIdx.x; u < total; u += gridDim.x) {
        const OpItem& it = find_item(one, items, nitems, u);
        const uint64_t c = u - it.begin;
        {
            const double2* pp = reinterpret_cast<const double2*>(it.partial);
            double m[4] = {0, 0, 0, 0};
            for (uint64_t i = threadIdx.x; i < it.count; i += kDiffChunkThreads) {
                const double2 a = pp[2 * i], b = pp[2 * i + 1];
                acc_merge(m[0], m[1], a.x, a.y);
                acc_merge(m[2], m[3], b.x, b.y);
            }
            diff_block_merge(m, sh);
            if (threadIdx.x == 0) {
                const int nbits = it.chunk_bits + (63 - __builtin_clzll(it.count));
                const double inv = ldexp(1.0, -nbits);   // 2^-n is exact: mean = sum * 2^-n
                mean[0] = 2.0 * ((m[0] + m[1]) * inv);
                mean[1] = 2.0 * ((m[2] + m[3]) * inv);
            }
            __syncthreads();
        }
        const double ar = mean[0], ai = mean[1];
        const uint64_t len = 1ull << it.chunk_bits;
        const uint64_t g0 = c << it.chunk_bits;
        T* re = static_cast<T*>(it.re) + g0;
        T* im = static_cast<T*>(it.im) + g0;
        const bool next = (it.flags & kChainNext) != 0;
        const bool flip = (it.flags & kChainFlip) != 0;
        const int variant = it.variant;
        const uint64_t fmask = it.fmask, fvalue = it.fvalue;
        double v[4] = {0, 0, 0, 0};
        if (it.chunk_bits >= kDiffVecMinBits) {
            V* rv = reinterpret_cast<V*>(re);
            V* iv = reinterpret_cast<V*>(im);
            const uint64_t nv = len / E;
            // four vectors of re and im in flight per step (loaded before any
            // store: re and im may not be assumed apart), updated in order
            constexpr int U = 4;
            for (uint64_t k0 = threadIdx.x; k0 < nv; k0 += U * kDiffChunkThreads) {
                V xr[U], xi[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint64_t k = k0 + static_cast<uint64_t>(u) * kDiffChunkThreads;
                    if (k < nv) {
                        xr[u] = rv[k];
                        xi[u] = iv[k];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint64_t k = k0 + static_cast<uint64_t>(u) * kDiffChunkThreads;
                    if (k < nv) {
                        T* a = reinterpret_cast<T*>(&xr[u]);
                        T* b = reinterpret_cast<T*>(&xi[u]);
#pragma unroll
                        for (int j = 0; j < E; ++j)
                            diff_update(a[j], b[j], ar, ai, variant, flip, g0 + k * E + j, fmask, fvalue, next, v);
                        rv[k] = xr[u];
                        iv[k] = xi[u];
                    }
                }
            }
        } else {
            for (uint64_t i = threadIdx.x; i < len; i += kDiffChunkThreads)
                diff_update(re[i], im[i], ar, ai, variant, flip, g0 + i, fmask, fvalue, next, v);
        }
        if (next) diff_store_chunk(v, it.partial_next, c, sh);
        __syncthreads();   // mean[] and sh[] are rewritten by the next chunk
    }
}

// DIFFUSE over t <= 12 bits `tmask`: one workgroup per block of the other
// bits; the block's 2^t amplitudes stay in registers between the
// compensated sum (fixed tree: deterministic) and the update.
constexpr int kDiffThreads = 256;
template <typename T>
__global__ __launch_bounds__(kDiffThreads) void diffuse_block_kernel(T* __restrict__ re, T* __restrict__ im, int t,
                                                                     uint64_t tmask, uint64_t others,
                                                                     uint64_t nblocks, int variant) {
    constexpr int kMaxE = (1 << kDiffuseMaxBits) / kDiffThreads;
    __shared__ double sh[2][2][kDiffThreads];
    const uint32_t count = 1u << t;
    const uint32_t tid = threadIdx.x;
    const double inv = 1.0 / static_cast<double>(count);
    for (uint64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        const uint64_t base = deposit(blk, others);
        T xr[kMaxE], xi[kMaxE];
        double sr = 0, cr = 0, si = 0, ci = 0;
#pragma unroll
        for (int r = 0; r < kMaxE; ++r) {
            const uint32_t e = tid + r * kDiffThreads;
            if (e < count) {
                const uint64_t idx = base | deposit(e, tmask);
                xr[r] = re[idx];
                xi[r] = im[idx];
                two_sum_add(sr, cr, static_cast<double>(xr[r]));
                two_sum_add(si, ci, static_cast<double>(xi[r]));
            }
        }
        sh[0][0][tid] = sr; sh[0][1][tid] = cr;
        sh[1][0][tid] = si; sh[1][1][tid] = ci;
        __syncthreads();
        for (int w = kDiffThreads / 2; w > 0; w >>= 1) {
            if (tid < static_cast<uint32_t>(w)) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    double ss = sh[q][0][tid], cc = sh[q][1][tid];
                    acc_merge(ss, cc, sh[q][0][tid + w], sh[q][1][tid + w]);
                    sh[q][0][tid] = ss;
                    sh[q][1][tid] = cc;
                }
            }
            __syncthreads();
        }
        const double ar = 2.0 * ((sh[0][0][0] + sh[0][1][0]) * inv);
        const double ai = 2.0 * ((sh[1][0][0] + sh[1][1][0]) * inv);
        __syncthreads();  // the totals are read before the next block reuses sh
#pragma unroll
        for (int r = 0; r < kMaxE; ++r) {
            const uint32_t e = tid + r * kDiffThreads;
            if (e < count) {
                const uint64_t idx = base | deposit(e, tmask);
                const double vr = xr[r], vi = xi[r];
                if (variant == 0) {
                    re[idx] = static_cast<T>(vr - ar);
                    im[idx] = static_cast<T>(vi - ai);
                } else {
                    re[idx] = static_cast<T>(ar - vr);
                    im[idx] = static_cast<T>(-ai - vi);
                }
            }
        }
    }
}

// DIFFUSE over t > kDiffuseMaxBits bits of a partial mask (apply_grover_diffuser
// takes any target list).  The amplitudes of one block of the other bits are
// enumerated in mask order: thread j starts at deposit(j, tmask) and steps by
// deposit(256, tmask) with the masked add ((d | ~tmask) + step) & tmask.
__device__ __forceinline__ uint64_t mask_add(uint64_t d, uint64_t step, uint64_t mask) {
    return ((d | ~mask) + step) & mask;
}

// Many blocks (more than 2^14): one workgroup per block sums its 2^t
// amplitudes (TwoSum per thread, the fixed merge of diff_block_merge), then
// updates them; the second sweep re-reads the block.
template <typename T>
__global__ __launch_bounds__(kDiffChunkThreads) void diffuse_wide_block_kernel(T* __restrict__ re, T* __restrict__ im,
                                                                              int t, uint64_t tmask, uint64_t others,
                                                                              uint64_t nblocks, int variant) {
    __shared__ double sh[kDiffWaves][4];
    __shared__ double mean[2];
    const uint64_t count = 1ull << t;
    const uint64_t first = deposit(threadIdx.x, tmask);
    const uint64_t step = deposit(kDiffChunkThreads, tmask);
    for (uint64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
        const uint64_t base = deposit(blk, others);
        double v[4] = {0, 0, 0, 0};
        uint64_t d = first;
        for (uint64_t e = threadIdx.x; e < count; e += kDiffChunkThreads) {
            two_sum_add(v[0], v[1], static_cast<double>(re[base | d]));
            two_sum_add(v[2], v[3], static_cast<double>(im[base | d]));
            d = mask_add(d, step, tmask);
        }
        diff_block_merge(v, sh);
        if (threadIdx.x == 0) {
            const double inv = ldexp(1.0, -t);
            mean[0] = 2.0 * ((v[0] + v[1]) * inv);
            mean[1] = 2.0 * ((v[2] + v[3]) * inv);
        }
        __syncthreads();
        const double ar = mean[0], ai = mean[1];
        double unused[4];
        d = first;
        for (uint64_t e = threadIdx.x; e < count; e += kDiffChunkThreads) {
            diff_update(re[base | d], im[base | d], ar, ai, variant, false, 0, 0, 0, false, unused);
            d = mask_add(d, step, tmask);
        }
        __syncthreads();   // mean[] is rewritten for the next block
    }
}

// Few blocks: block b's 2^t amplitudes split into nch = 2^lch chunks of 2^cb
// (chunk c = the amplitudes whose top lch mask bits read c); pass 1 writes
// one compensated pair per (block, chunk) unit, pass 2 merges the block's
// nch pairs in a fixed order in every unit and updates the unit's chunk.
template <typename T>
__global__ __launch_bounds__(kDiffChunkThreads) void diffuse_wide_partial_kernel(
    const T* __restrict__ re, const T* __restrict__ im, int cb, int lch, uint64_t tmask, uint64_t others,
    uint64_t units, double* __restrict__ partial) {
    __shared__ double sh[kDiffWaves][4];
    const uint64_t len = 1ull << cb;
    const uint64_t first = deposit(threadIdx.x, tmask);
    const uint64_t step = deposit(kDiffChunkThreads, tmask);
    for (uint64_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint64_t b = u >> lch, c = u & ((1ull << lch) - 1);
        const uint64_t base = deposit(b, others) | deposit(c << cb, tmask);
        double v[4] = {0, 0, 0, 0};
        uint64_t d = first;
        for (uint64_t e = threadIdx.x; e < len; e += kDiffChunkThreads) {
            two_sum_add(v[0], v[1], static_cast<double>(re[base | d]));
            two_sum_add(v[2], v[3], static_cast<double>(im[base | d]));
            d = mask_add(d, step, tmask);
        }
        diff_store_chunk(v, partial, u, sh);
    }
}

template <typename T>
__global__ __launch_bounds__(kDiffChunkThreads) void diffuse_wide_apply_kernel(
    T* __restrict__ re, T* __restrict__ im, int t, int cb, int lch, uint64_t tmask, uint64_t others, uint64_t units,
    const double* __restrict__ partial, int variant) {
    __shared__ double sh[kDiffWaves][4];
    __shared__ double mean[2];
    const uint64_t len = 1ull << cb;
    const uint64_t nch = 1ull << lch;
    const uint64_t first = deposit(threadIdx.x, tmask);
    const uint64_t step = deposit(kDiffChunkThreads, tmask);
    for (uint64_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint64_t b = u >> lch, c = u & (nch - 1);
        {
            const double2* pp = reinterpret_cast<const double2*>(partial) + 2 * (b << lch);
            double m[4] = {0, 0, 0, 0};
            for (uint64_t i = threadIdx.x; i < nch; i += kDiffChunkThreads) {
                const double2 x = pp[2 * i], y = pp[2 * i + 1];
                acc_merge(m[0], m[1], x.x, x.y);
                acc_merge(m[2], m[3], y.x, y.y);
            }
            diff_block_merge(m, sh);
            if (threadIdx.x == 0) {
                const double inv = ldexp(1.0, -t);
                mean[0] = 2.0 * ((m[0] + m[1]) * inv);
                mean[1] = 2.0 * ((m[2] + m[3]) * inv);
            }
            __syncthreads();
        }
        const double ar = mean[0], ai = mean[1];
        const uint64_t base = deposit(b, others) | deposit(c << cb, tmask);
        double unused[4];
        uint64_t d = first;
        for (uint64_t e = threadIdx.x; e < len; e += kDiffChunkThreads) {
            diff_update(re[base | d], im[base | d], ar, ai, variant, false, 0, 0, 0, false, unused);
            d = mask_add(d, step, tmask);
        }
        __syncthreads();   // mean[] is rewritten by the next unit
    }
}

constexpr int kMargThreads = 64;
constexpr int kMargMaxBits = 6;

// Marginal over m target bits: each thread owns 2^m compensated bins in LDS.
template <typename T>
__global__ __launch_bounds__(kMargThreads) void marginal_kernel(const T* __restrict__ re, const T* __restrict__ im,
                                                                uint64_t size, int m, int t0, int t1, int t2, int t3,
                                                                int t4, int t5, double* __restrict__ partial) {
    __shared__ double bins[kMargThreads][1 << kMargMaxBits][2];
    const int nb = 1 << m;
    const int t[6] = {t0, t1, t2, t3, t4, t5};
    for (int b = 0; b < nb; ++b) { bins[threadIdx.x][b][0] = 0; bins[threadIdx.x][b][1] = 0; }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kMargThreads;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kMargThreads + threadIdx.x; i < size; i += stride) {
        int b = 0;
        for (int q = 0; q < m; ++q) b |= static_cast<int>((i >> t[q]) & 1ull) << q;
        const double ar = re[i], ai = im[i];
        double s = bins[threadIdx.x][b][0], c = bins[threadIdx.x][b][1];
        two_sum_add(s, c, ar * ar + ai * ai);
        bins[threadIdx.x][b][0] = s;
        bins[threadIdx.x][b][1] = c;
    }
    __syncthreads();
    // fixed-order merge across threads, one bin per thread
    for (int b = threadIdx.x; b < nb; b += kMargThreads) {
        double s = 0, c = 0;
        for (int th = 0; th < kMargThreads; ++th) acc_merge(s, c, bins[th][b][0], bins[th][b][1]);
        partial[2 * (b * gridDim.x + blockIdx.x)] = s;
        partial[2 * (b * gridDim.x + blockIdx.x) + 1] = c;
    }
}

// Blocks of 256 threads for a grid-stride kernel: one per `per_block` units
// of work (the full grid of the streaming launches), or at most `cap` when
// one is given.
inline unsigned grid_for(uint64_t work, unsigned per_block = 256, unsigned cap = 0) {
    uint64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (cap) return static_cast<unsigned>(g > cap ? cap : g);
    return full_grid(g, 16384, 256);
}

template <typename T>
hipError_t launch_gate_t(T* re, T* im, const GateLaunch& G, hipStream_t st) {
    hipLaunchKernelGGL(gate_kernel<T>, dim3(full_grid((G.count + 255) / 256, 16384, 256)), dim3(256), 0, st, re, im, G);
    return hipGetLastError();
}

// Persistent grid of 4096 blocks (W = 32: 34 KiB of LDS for complex128,
// 17 KiB for complex64).  Measured at 30 qubits: complex128 7.26 ms, complex64
// 4.02 ms, against 9.36 / 5.33 ms for re and im in separate phases with one
// block per tile pair in index order.
constexpr unsigned kBitrevBlocks = 4096;

template <typename T>
hipError_t launch_bitrev_t(T* re, T* im, int n, uint64_t mask, hipStream_t st) {
    const uint64_t size = 1ull << n;
    if (mask == 0 && n >= 10) {
        const uint64_t tiles = 1ull << (n - 10);
        const unsigned blocks = full_grid(tiles, kBitrevBlocks, 256);
        hipLaunchKernelGGL((bitrev_joint_kernel<T, 5>), dim3(blocks), dim3(256), 0, st, re, im, n);
    } else {
        if (mask == 0) mask = size - 1;
        hipLaunchKernelGGL(bitrev_simple_kernel<T>, dim3(grid_for(size)), dim3(256), 0, st, re, im, size, mask, n);
    }
    return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------- measurement
// Outcome probabilities P(o) = sum of |a_i|^2 over i = dep(o, free bits) |
// rest | fval, the rest enumerating the bits of `rmask` (neither free nor
// forced).  Block b takes the rest values [b * rpb, (b + 1) * rpb).  With
// 2^f >= NT a thread owns outcomes tid + NT k (k < NO); with 2^f < NT the
// NT / 2^f threads of one outcome interleave over the rest values and are
// merged through LDS in thread order.  Per-thread TwoSum, fixed assignment and
// order: deterministic.  Rest indices advance by masked addition:
// dep(r + d) = ((dep(r) | ~rmask) + dep(d)) & rmask.
constexpr int kOutThreads = 256;

__host__ __device__ inline uint64_t deposit64(uint64_t x, uint64_t mask) {
    uint64_t out = 0;
    while (mask) {
        const uint64_t low = mask & (~mask + 1);
        if (x & 1) out |= low;
        x >>= 1;
        mask ^= low;
    }
    return out;
}

template <typename T, int NO>
__global__ __launch_bounds__(kOutThreads) void outcome_partials_kernel(const T* __restrict__ re,
                                                                       const T* __restrict__ im, OutcomeSpec sp,
                                                                       double* __restrict__ part) {
    __shared__ double s_part[kOutThreads];
    const uint32_t nout = 1u << sp.f;
    const uint32_t tid = threadIdx.x;
    const bool wide = nout >= static_cast<uint32_t>(kOutThreads);
    const uint32_t o0 = wide ? tid : tid % nout;
    const uint64_t rsub = wide ? 0 : tid / nout;
    const uint64_t rstep = wide ? 1 : kOutThreads / nout;
    uint64_t dep_o[NO];
#pragma unroll
    for (int k = 0; k < NO; ++k) {
        const uint32_t o = o0 + static_cast<uint32_t>(k * kOutThreads);
        uint64_t d = 0;
        for (int q = 0; q < sp.f; ++q) d |= static_cast<uint64_t>((o >> q) & 1u) << sp.fpos[q];
        dep_o[k] = d;
    }
    double s[NO], c[NO];
#pragma unroll
    for (int k = 0; k < NO; ++k) { s[k] = 0.0; c[k] = 0.0; }
    const uint64_t r_begin = static_cast<uint64_t>(blockIdx.x) * sp.rest_per_block + rsub;
    const uint64_t r_end = static_cast<uint64_t>(blockIdx.x + 1) * sp.rest_per_block;
    uint64_t x = deposit64(r_begin, sp.rmask);
    const uint64_t dstep = deposit64(rstep, sp.rmask);
    for (uint64_t r = r_begin; r < r_end; r += rstep) {
        const uint64_t base = x | sp.fval;
#pragma unroll
        for (int k = 0; k < NO; ++k) {
            const uint64_t i = base | dep_o[k];
            const double a = static_cast<double>(re[i]), b = static_cast<double>(im[i]);
            two_sum_add(s[k], c[k], a * a + b * b);
        }
        x = ((x | ~sp.rmask) + dstep) & sp.rmask;
    }
    if (wide) {
#pragma unroll
        for (int k = 0; k < NO; ++k)
            part[static_cast<uint64_t>(blockIdx.x) * nout + o0 + static_cast<uint32_t>(k * kOutThreads)] = s[k] + c[k];
        return;
    }
    s_part[tid] = s[0] + c[0];
    __syncthreads();
    if (tid < nout) {
        double ss = 0.0, cc = 0.0;
        for (uint32_t j = 0; j < rstep; ++j) two_sum_add(ss, cc, s_part[tid + j * nout]);
        part[static_cast<uint64_t>(blockIdx.x) * nout + tid] = ss + cc;
    }
}

// out[o] = sum over blocks of part[b][o] (TwoSum, block order).
__global__ void outcome_merge_kernel(const double* __restrict__ part, int nblocks, uint32_t nout,
                                     double* __restrict__ out) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nblocks; ++b) two_sum_add(s, c, part[static_cast<uint64_t>(b) * nout + o]);
    out[o] = s + c;
}

// Collapse: amplitudes with (i & tmask) == tval times inv, the others 0
// (apply_measure's last loop, state.mojo:535-553).
template <typename T>
__global__ void collapse_kernel(T* __restrict__ re, T* __restrict__ im, uint64_t size, uint64_t tmask, uint64_t tval,
                                T inv) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < size; i += stride) {
        if ((i & tmask) == tval) {
            re[i] = re[i] * inv;
            im[i] = im[i] * inv;
        } else {
            re[i] = T(0);
            im[i] = T(0);
        }
    }
}

// ---------------------------------------------------------------- wrappers
hipError_t launch_tile_pass(int dtype, void* re, void* im, const TilePass& P, const TileSweep* d_sweeps,
                            const TileInstr* d_instrs, hipStream_t stream) {
    if (dtype == BFLY_C64)
        return launch_tile_t<float>(static_cast<float*>(re), static_cast<float*>(im), P, d_sweeps, d_instrs, stream);
    return launch_tile_t<double>(static_cast<double*>(re), static_cast<double*>(im), P, d_sweeps, d_instrs, stream);
}

hipError_t prepare_tile_kernels(int dtype, int k, int slots) {
    TilePass P{};
    P.k = k;
    P.n = k;
    P.slots = slots;
    hipError_t e = launch_tile_pass(dtype, nullptr, nullptr, P, nullptr, nullptr, nullptr);
    if (e != hipSuccess) return e;
    return launch_tile_batch(dtype, k, slots, nullptr, 0, 0, nullptr);
}

hipError_t launch_tile_batch(int dtype, int k, int slots, const BatchItem* d_items, int nitems, uint64_t total_tiles,
                             hipStream_t stream) {
    if (dtype == BFLY_C64) return launch_batch_t<float>(k, slots, d_items, nitems, total_tiles, stream);
    return launch_batch_t<double>(k, slots, d_items, nitems, total_tiles, stream);
}

hipError_t launch_block_pass(int dtype, void* re, void* im, const TilePass& P, const BlockPass& B, const float* mat,
                             const double* tab, const double* norm2, hipStream_t stream) {
    if (dtype == BFLY_C128) {
        double* r = static_cast<double*>(re);
        double* i = static_cast<double*>(im);
        const double* m = reinterpret_cast<const double*>(mat);
        switch (P.k) {
            case 9: return launch_block64_k<9>(r, i, P, B, m, tab, norm2, stream);
            case 10: return launch_block64_k<10>(r, i, P, B, m, tab, norm2, stream);
            case 11:
                if (P.nouter <= kWs64OuterMax) return launch_block_ws64(r, i, P, B, m, tab, norm2, stream);
                return launch_block64_k<11>(r, i, P, B, m, tab, norm2, stream);
        }
        return hipErrorInvalidValue;
    }
    float* r = static_cast<float*>(re);
    float* i = static_cast<float*>(im);
    switch (P.k) {
        case 9: return launch_block_k<9>(r, i, P, B, mat, tab, norm2, stream);
        case 10: return launch_block_k<10>(r, i, P, B, mat, tab, norm2, stream);
        case 11: return launch_block_k<11>(r, i, P, B, mat, tab, norm2, stream);
        case 12:
            if (P.nouter <= kWsOuterMax) return launch_block_ws(r, i, P, B, mat, tab, norm2, stream);
            return launch_block_k<12>(r, i, P, B, mat, tab, norm2, stream);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_gate(int dtype, void* re, void* im, const GateLaunch& G, hipStream_t stream) {
    if (dtype == BFLY_C64) return launch_gate_t<float>(static_cast<float*>(re), static_cast<float*>(im), G, stream);
    return launch_gate_t<double>(static_cast<double*>(re), static_cast<double*>(im), G, stream);
}

hipError_t launch_bitrev(int dtype, void* re, void* im, int n, uint64_t mask, hipStream_t stream) {
    if (dtype == BFLY_C64) return launch_bitrev_t<float>(static_cast<float*>(re), static_cast<float*>(im), n, mask, stream);
    return launch_bitrev_t<double>(static_cast<double*>(re), static_cast<double*>(im), n, mask, stream);
}

hipError_t launch_swap(int dtype, void* re, void* im, int n, int a, int b, hipStream_t stream) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const uint64_t count = 1ull << (n - 2);
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(swap_kernel<float>, dim3(grid_for(count)), dim3(256), 0, stream, static_cast<float*>(re),
                           static_cast<float*>(im), count, lo, hi, a, b);
    else
        hipLaunchKernelGGL(swap_kernel<double>, dim3(grid_for(count)), dim3(256), 0, stream, static_cast<double*>(re),
                           static_cast<double*>(im), count, lo, hi, a, b);
    return hipGetLastError();
}

// Dense 2^M blocks, M = 4..6, on MFMA: every wave takes 16 groups (the N
// dimension) at a time and computes, with A = [Ur, -Ui] (D x 2D, D = 2^M),
//   [re'] = A [re; im]   and   [im'] = A [im; -re]
// over the K = 2D inputs of each group.  A is staged in LDS once per
// workgroup as At[k][row] (k-major); the lane's B operand (its group's inputs
// k = 4 s + lane / 16) is loaded from HBM into registers, all 2D of them
// before any output of the wave's groups is written, so the update is in
// place.  complex128: v_mfma_f64_16x16x4f64 (FP64 products and sums; lane l
// holds output rows l / 16 + 4 v of a 16-row block); complex64:
// v_mfma_f32_16x16x4f32 (FP32 products and sums, the reference's FLOAT32
// arithmetic; lane l holds rows 4 (l / 16) + v).  Groups past the end load
// zeros and store nothing: an MFMA column depends on its own B column only.
constexpr int kUMfmaThreads = 256;

template <typename T, int M>
__global__ __launch_bounds__(kUMfmaThreads) void unitary_mfma_kernel(T* __restrict__ re, T* __restrict__ im,
                                                                     const UnitaryArgs A,
                                                                     const double* __restrict__ u) {
    constexpr int D = 1 << M;
    constexpr int NS = D / 4;    // K steps per half (re part, im part)
    constexpr int NR = D / 16;   // 16-row output blocks
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* At = reinterpret_cast<T*>(lds_raw);   // At[k * D + row], k < 2D
    for (int i = threadIdx.x; i < 2 * D * D; i += kUMfmaThreads) {
        const int k = i / D, row = i % D;
        At[i] = k < D ? static_cast<T>(u[2 * (row * D + k)]) : static_cast<T>(-u[2 * (row * D + (k - D)) + 1]);
    }
    __syncthreads();
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int g = lane >> 4, c = lane & 15;
    constexpr bool kF64 = sizeof(T) == 8;
    typedef T acc_t __attribute__((ext_vector_type(4)));
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * (kUMfmaThreads / 4);
    for (uint64_t g0 = (static_cast<uint64_t>(blockIdx.x) * (kUMfmaThreads / 64) + wave) * 16; g0 < A.ngroups;
         g0 += step) {
        const uint64_t grp = g0 + static_cast<uint64_t>(c);
        const bool live = grp < A.ngroups;
        const uint64_t base = live ? unitary_base(grp, A) : 0;
        T xr[NS], xi[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const uint64_t idx = base + (static_cast<uint64_t>(4 * s + g) << A.target);
            xr[s] = live ? re[idx] : T(0);
            xi[s] = live ? im[idx] : T(0);
        }
        // one 16-row output block at a time keeps 16 accumulator values per
        // lane live (beside the 2D inputs), not all D rows
#pragma unroll 1
        for (int r = 0; r < NR; ++r) {
            acc_t acc_r = acc_t{T(0), T(0), T(0), T(0)};
            acc_t acc_i = acc_t{T(0), T(0), T(0), T(0)};
            const T* at = At + 16 * r + c;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const T a_re = at[(4 * s + g) * D];         // Ur[row][4 s + g]
                const T a_im = at[(D + 4 * s + g) * D];     // -Ui[row][4 s + g]
                if constexpr (kF64) {
                    acc_r = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re, xr[s], acc_r, 0, 0, 0);
                    acc_r = __builtin_amdgcn_mfma_f64_16x16x4f64(a_im, xi[s], acc_r, 0, 0, 0);
                    acc_i = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re, xi[s], acc_i, 0, 0, 0);
                    acc_i = __builtin_amdgcn_mfma_f64_16x16x4f64(a_im, -xr[s], acc_i, 0, 0, 0);
                } else {
                    acc_r = __builtin_amdgcn_mfma_f32_16x16x4f32(a_re, xr[s], acc_r, 0, 0, 0);
                    acc_r = __builtin_amdgcn_mfma_f32_16x16x4f32(a_im, xi[s], acc_r, 0, 0, 0);
                    acc_i = __builtin_amdgcn_mfma_f32_16x16x4f32(a_re, xi[s], acc_i, 0, 0, 0);
                    acc_i = __builtin_amdgcn_mfma_f32_16x16x4f32(a_im, -xr[s], acc_i, 0, 0, 0);
                }
            }
            // every input of the wave's groups is in registers already: the
            // rows of this block can be stored before the next block runs
            if (live) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = 16 * r + (kF64 ? g + 4 * v : 4 * g + v);
                    const uint64_t idx = base + (static_cast<uint64_t>(row) << A.target);
                    re[idx] = acc_r[v];
                    im[idx] = acc_i[v];
                }
            }
        }
    }
}

template <typename T, int M>
hipError_t launch_unitary_mfma_t(T* re, T* im, const UnitaryArgs& A, const double* d_u, hipStream_t stream) {
    constexpr int D = 1 << M;
    const size_t lds = 2ull * D * D * sizeof(T);
    const uint64_t waves = (A.ngroups + 15) / 16;
    const uint64_t blocks = (waves + 3) / 4;
    // Every block stages the matrix in LDS: 4-qubit blocks (4 KiB in
    // complex128) take one block per 4 waves of groups; 5 and 6 qubits keep
    // a persistent grid that loads it once per block.  Measured at 28
    // qubits (complex128 / complex64): m = 4 1.49 / 0.95 ms full against
    // 1.79 / 1.23 ms persistent; m = 5 1.90 / 1.37 against 1.75 / 1.26;
    // m = 6 3.45 / 2.40 against 2.40 / 1.45.
    const unsigned grid =
        M <= 4 ? full_grid(blocks, 4096, kUMfmaThreads) : static_cast<unsigned>(blocks < 4096 ? blocks : 4096);
    hipLaunchKernelGGL((unitary_mfma_kernel<T, M>), dim3(grid), dim3(kUMfmaThreads), lds, stream, re, im, A, d_u);
    return hipGetLastError();
}

hipError_t launch_unitary(int dtype, void* re, void* im, int n, int target, int m, uint64_t controls,
                          const double* d_u, hipStream_t stream) {
    UnitaryArgs A{};
    A.controls = controls;
    A.target = target;
    A.m = m;
    const uint64_t bits = (((1ull << m) - 1) << target) | controls;
    for (int b = 0; b < n; ++b)
        if ((bits >> b) & 1) A.pos[A.npos++] = b;
    A.ngroups = 1ull << (n - A.npos);
    if (m >= 4 && m <= 6) {
        if (dtype == BFLY_C64) {
            float* r = static_cast<float*>(re);
            float* i = static_cast<float*>(im);
            return m == 4 ? launch_unitary_mfma_t<float, 4>(r, i, A, d_u, stream)
                 : m == 5 ? launch_unitary_mfma_t<float, 5>(r, i, A, d_u, stream)
                          : launch_unitary_mfma_t<float, 6>(r, i, A, d_u, stream);
        }
        double* r = static_cast<double*>(re);
        double* i = static_cast<double*>(im);
        return m == 4 ? launch_unitary_mfma_t<double, 4>(r, i, A, d_u, stream)
             : m == 5 ? launch_unitary_mfma_t<double, 5>(r, i, A, d_u, stream)
                      : launch_unitary_mfma_t<double, 6>(r, i, A, d_u, stream);
    }
    const int dim = 1 << m;
    const int G = dim >= 256 ? 1 : 256 / dim;
    const size_t lds = 2ull * G * dim * (dtype == BFLY_C64 ? sizeof(float) : sizeof(double));
    const unsigned grid = grid_for(A.ngroups, static_cast<unsigned>(G), 65535);
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(unitary_kernel<float>, dim3(grid), dim3(256), lds, stream, static_cast<float*>(re),
                           static_cast<float*>(im), A, d_u);
    else
        hipLaunchKernelGGL(unitary_kernel<double>, dim3(grid), dim3(256), lds, stream, static_cast<double*>(re),
                           static_cast<double*>(im), A, d_u);
    return hipGetLastError();
}

// Clears n 64-bit words (the tile counters of work-queue passes) with a
// kernel, so that a captured plan execution records no memset node.
__global__ __launch_bounds__(64) void zero_words_kernel(unsigned long long* __restrict__ p, int n) {
    const int i = static_cast<int>(blockIdx.x * 64u + threadIdx.x);
    if (i < n) p[i] = 0ull;
}

hipError_t launch_zero_words(unsigned long long* p, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_words_kernel, dim3(static_cast<unsigned>((n + 63) / 64)), dim3(64), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_reset(int dtype, void* re, void* im, uint64_t size, hipStream_t stream) {
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(reset_kernel<float>, dim3(grid_for(size)), dim3(256), 0, stream, static_cast<float*>(re),
                           static_cast<float*>(im), size);
    else
        hipLaunchKernelGGL(reset_kernel<double>, dim3(grid_for(size)), dim3(256), 0, stream, static_cast<double*>(re),
                           static_cast<double*>(im), size);
    return hipGetLastError();
}

hipError_t launch_random_fill(int dtype, void* re, void* im, uint64_t size, uint64_t seed, hipStream_t stream) {
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(random_fill_kernel<float>, dim3(grid_for(size)), dim3(256), 0, stream,
                           static_cast<float*>(re), static_cast<float*>(im), size, seed);
    else
        hipLaunchKernelGGL(random_fill_kernel<double>, dim3(grid_for(size)), dim3(256), 0, stream,
                           static_cast<double*>(re), static_cast<double*>(im), size, seed);
    return hipGetLastError();
}

// Blocks for the streaming kernels: one vector per thread per unrolled step,
// at least one block.
unsigned stream_grid(uint64_t nvec) {
    const uint64_t per_block = 256ull * kStreamUnroll;
    const uint64_t b = (nvec + per_block - 1) / per_block;
    return full_grid(b < 1 ? 1 : b, 32768, 256);
}

hipError_t launch_scale(int dtype, void* re, void* im, uint64_t size, double scale, hipStream_t stream) {
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(scale_kernel<float>, dim3(stream_grid(size / 4)), dim3(256), 0, stream,
                           static_cast<float*>(re), static_cast<float*>(im), size, scale);
    else
        hipLaunchKernelGGL(scale_kernel<double>, dim3(stream_grid(size / 2)), dim3(256), 0, stream,
                           static_cast<double*>(re), static_cast<double*>(im), size, scale);
    return hipGetLastError();
}

hipError_t launch_convert(int src_dtype, const void* sre, const void* sim, void* dre, void* dim, uint64_t size,
                          hipStream_t stream) {
    if (src_dtype == BFLY_C64)
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(size)), dim3(256), 0, stream,
                           static_cast<const float*>(sre), static_cast<const float*>(sim), static_cast<double*>(dre),
                           static_cast<double*>(dim), size);
    else
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(size)), dim3(256), 0, stream,
                           static_cast<const double*>(sre), static_cast<const double*>(sim), static_cast<float*>(dre),
                           static_cast<float*>(dim), size);
    return hipGetLastError();
}

hipError_t launch_normalize(int dtype, void* re, void* im, uint64_t size, const double* d_norm2, hipStream_t stream) {
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(normalize_kernel<float>, dim3(grid_for(size)), dim3(256), 0, stream, static_cast<float*>(re),
                           static_cast<float*>(im), size, d_norm2);
    else
        hipLaunchKernelGGL(normalize_kernel<double>, dim3(grid_for(size)), dim3(256), 0, stream,
                           static_cast<double*>(re), static_cast<double*>(im), size, d_norm2);
    return hipGetLastError();
}

int outcome_blocks(const OutcomeSpec& sp, uint64_t rest_count) {
    // at most 2^20 partial doubles; rest values split evenly (powers of two)
    uint64_t nb = 1024;
    while (nb > 1 && (nb << sp.f) > (1ull << 20)) nb >>= 1;
    while (nb > 1 && nb > rest_count) nb >>= 1;
    return static_cast<int>(nb);
}

template <typename T>
hipError_t launch_outcome_t(const T* re, const T* im, const OutcomeSpec& sp, int nblocks, double* part, double* out,
                            hipStream_t st) {
    const uint32_t nout = 1u << sp.f;
    const int no = nout > static_cast<uint32_t>(kOutThreads) ? static_cast<int>(nout / kOutThreads) : 1;
    switch (no) {
        case 1: hipLaunchKernelGGL((outcome_partials_kernel<T, 1>), dim3(nblocks), dim3(kOutThreads), 0, st, re, im, sp, part); break;
        case 2: hipLaunchKernelGGL((outcome_partials_kernel<T, 2>), dim3(nblocks), dim3(kOutThreads), 0, st, re, im, sp, part); break;
        case 4: hipLaunchKernelGGL((outcome_partials_kernel<T, 4>), dim3(nblocks), dim3(kOutThreads), 0, st, re, im, sp, part); break;
        case 8: hipLaunchKernelGGL((outcome_partials_kernel<T, 8>), dim3(nblocks), dim3(kOutThreads), 0, st, re, im, sp, part); break;
        case 16: hipLaunchKernelGGL((outcome_partials_kernel<T, 16>), dim3(nblocks), dim3(kOutThreads), 0, st, re, im, sp, part); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(outcome_merge_kernel, dim3((nout + 255) / 256), dim3(256), 0, st, part, nblocks, nout, out);
    return hipGetLastError();
}

// Grid (outcome blocks, R rest chunks): chunk y sums rest values
// [y*nrest/R, (y+1)*nrest/R) of outcome o into part[y][o]; with R == 1 part is
// the result.
template <typename T>
__global__ __launch_bounds__(256) void outcome_direct_kernel(const T* __restrict__ re, const T* __restrict__ im,
                                                             const OutcomeDirect sp, uint64_t per_chunk,
                                                             double* __restrict__ part) {
    const uint64_t nout = 1ull << sp.f;
    const uint64_t o = static_cast<uint64_t>(blockIdx.x) * 256u + threadIdx.x;
    if (o >= nout) return;
    uint64_t base = sp.fval;
    for (int j = 0; j < sp.f; ++j) base |= ((o >> j) & 1ull) << sp.fpos[j];
    // first rest value of this chunk, deposited on rmask
    const uint64_t r0 = static_cast<uint64_t>(blockIdx.y) * per_chunk;
    uint64_t dep = 0, m = sp.rmask;
    for (int b = 0; m; ++b, m &= m - 1)
        if ((r0 >> b) & 1ull) dep |= m & (~m + 1ull);
    double s = 0.0, c = 0.0;
    for (uint64_t r = 0; r < per_chunk; ++r) {
        const uint64_t i = base | dep;
        const double a = static_cast<double>(re[i]), b = static_cast<double>(im[i]);
        two_sum_add(s, c, a * a + b * b);
        dep = ((dep | ~sp.rmask) + 1ull) & sp.rmask;   // next rest value
    }
    part[static_cast<uint64_t>(blockIdx.y) * nout + o] = s + c;
}

// out[o] = the R partial sums of outcome o, added in chunk order.
__global__ __launch_bounds__(256) void outcome_direct_combine_kernel(const double* __restrict__ part, uint64_t nout,
                                                                     int R, double* __restrict__ out) {
    const uint64_t o = static_cast<uint64_t>(blockIdx.x) * 256u + threadIdx.x;
    if (o >= nout) return;
    double s = 0.0, c = 0.0;
    for (int y = 0; y < R; ++y) two_sum_add(s, c, part[static_cast<uint64_t>(y) * nout + o]);
    out[o] = s + c;
}

int outcome_direct_chunks(const OutcomeDirect& sp) {
    const uint64_t gx = ((1ull << sp.f) + 255) / 256;
    int R = 1;
    while (gx * static_cast<uint64_t>(R) < 2048 && static_cast<uint64_t>(R) * 2 <= sp.nrest && R < 1024) R *= 2;
    return R;
}

hipError_t launch_outcome_direct(int dtype, const void* re, const void* im, const OutcomeDirect& sp, int R,
                                 double* part, double* out, hipStream_t st) {
    if (sp.f < 1 || sp.f > kOutcomeDirectMaxBits || R < 1 || sp.nrest % static_cast<uint64_t>(R))
        return hipErrorInvalidValue;
    const uint64_t nout = 1ull << sp.f;
    const unsigned gx = static_cast<unsigned>((nout + 255) / 256);
    const uint64_t per = sp.nrest / static_cast<uint64_t>(R);
    double* dst = R == 1 ? out : part;
    if (dtype == BFLY_C128)
        hipLaunchKernelGGL(outcome_direct_kernel<double>, dim3(gx, R), dim3(256), 0, st,
                           static_cast<const double*>(re), static_cast<const double*>(im), sp, per, dst);
    else
        hipLaunchKernelGGL(outcome_direct_kernel<float>, dim3(gx, R), dim3(256), 0, st,
                           static_cast<const float*>(re), static_cast<const float*>(im), sp, per, dst);
    if (R > 1) hipLaunchKernelGGL(outcome_direct_combine_kernel, dim3(gx), dim3(256), 0, st, part, nout, R, out);
    return hipGetLastError();
}

hipError_t launch_outcome_probs(int dtype, const void* re, const void* im, const OutcomeSpec& sp, int nblocks,
                                double* part, double* out, hipStream_t st) {
    if (sp.f < 0 || sp.f > kOutcomeMaxBits) return hipErrorInvalidValue;
    if (dtype == BFLY_C128)
        return launch_outcome_t(static_cast<const double*>(re), static_cast<const double*>(im), sp, nblocks, part, out, st);
    return launch_outcome_t(static_cast<const float*>(re), static_cast<const float*>(im), sp, nblocks, part, out, st);
}

hipError_t launch_collapse(int dtype, void* re, void* im, uint64_t size, uint64_t tmask, uint64_t tval, double inv,
                           hipStream_t st) {
    const unsigned grid = static_cast<unsigned>(std::min<uint64_t>((size + 255) / 256, 4096));
    if (dtype == BFLY_C128)
        hipLaunchKernelGGL(collapse_kernel<double>, dim3(grid), dim3(256), 0, st, static_cast<double*>(re),
                           static_cast<double*>(im), size, tmask, tval, inv);
    else
        hipLaunchKernelGGL(collapse_kernel<float>, dim3(grid), dim3(256), 0, st, static_cast<float*>(re),
                           static_cast<float*>(im), size, tmask, tval, static_cast<float>(inv));
    return hipGetLastError();
}

size_t reduce_scratch_doubles() {
    const size_t a = 2ull * 2 * kRedBlocks + 8;
    const size_t b = 2ull * (1 << kMargMaxBits) * kRedBlocks + (1 << kMargMaxBits);
    return a > b ? a : b;
}

namespace {
template <typename T, int MODE>
void launch_reduce_t(const void* are, const void* aim, const void* bre, const void* bim, uint64_t size, uint64_t mask,
                     uint64_t value, double* partial, unsigned nb, hipStream_t st) {
    hipLaunchKernelGGL((reduce_kernel<T, MODE>), dim3(nb), dim3(kRedThreads), 0, st, static_cast<const T*>(are),
                       static_cast<const T*>(aim), static_cast<const T*>(bre), static_cast<const T*>(bim), size, mask,
                       value, partial);
}
}  // namespace

unsigned reduce_blocks(uint64_t size) { return grid_for(size, kRedThreads * 4, kRedBlocks); }

hipError_t launch_reduce_into(int dtype, int mode, const void* are, const void* aim, const void* bre, const void* bim,
                              uint64_t size, uint64_t mask, uint64_t value, double* partial, double* out,
                              hipStream_t stream) {
    // the block count depends only on the size: same size -> same tree -> same bits
    const unsigned nb = reduce_blocks(size);
    const bool f = dtype == BFLY_C64;
    switch (mode) {
        case R_NORM2: f ? launch_reduce_t<float, R_NORM2>(are, aim, bre, bim, size, mask, value, partial, nb, stream)
                        : launch_reduce_t<double, R_NORM2>(are, aim, bre, bim, size, mask, value, partial, nb, stream); break;
        case R_PROB_MASK: f ? launch_reduce_t<float, R_PROB_MASK>(are, aim, bre, bim, size, mask, value, partial, nb, stream)
                            : launch_reduce_t<double, R_PROB_MASK>(are, aim, bre, bim, size, mask, value, partial, nb, stream); break;
        case R_DIFF2: f ? launch_reduce_t<float, R_DIFF2>(are, aim, bre, bim, size, mask, value, partial, nb, stream)
                        : launch_reduce_t<double, R_DIFF2>(are, aim, bre, bim, size, mask, value, partial, nb, stream); break;
        case R_INNER: f ? launch_reduce_t<float, R_INNER>(are, aim, bre, bim, size, mask, value, partial, nb, stream)
                        : launch_reduce_t<double, R_INNER>(are, aim, bre, bim, size, mask, value, partial, nb, stream); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nch = mode == R_INNER ? 2 : 1;
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kRedThreads), 0, stream, partial, static_cast<int>(nb), nch,
                       out);
    return hipGetLastError();
}

hipError_t launch_reduce(int dtype, int mode, const void* are, const void* aim, const void* bre, const void* bim,
                         uint64_t size, uint64_t mask, uint64_t value, double* scratch, hipStream_t stream) {
    return launch_reduce_into(dtype, mode, are, aim, bre, bim, size, mask, value, scratch + 8, scratch, stream);
}

OpItem flip_item(void* re, void* im, int n, uint64_t mask, uint64_t value) {
    OpItem it{};
    it.re = re;
    it.im = im;
    it.others = ((1ull << n) - 1) & ~mask;
    it.value = value;
    it.count = 1ull << (n - __builtin_popcountll(mask));
    return it;
}

// Chunks of at least 2^12 amplitudes (16 per thread) and at most 2^10
// chunks: every apply workgroup merges at most 32 KB of pairs, and two
// partial buffers fit the state's reduction scratch.
OpItem diffuse_item(void* re, void* im, int n, int variant, double* scratch, int32_t chain, uint64_t fmask,
                    uint64_t fvalue) {
    static_assert(2 * kPartialSpan <= 2ull * (1 << kMargMaxBits) * kRedBlocks + (1 << kMargMaxBits),
                  "two partial buffers fit the state's reduction scratch");
    constexpr int kMinChunkBits = 12, kMaxCountBits = 10;
    static_assert(4ull << kMaxCountBits <= kPartialSpan, "the pairs fit one partial buffer");
    OpItem it{};
    it.re = re;
    it.im = im;
    it.chunk_bits = n <= kMinChunkBits ? n : (n - kMaxCountBits > kMinChunkBits ? n - kMaxCountBits : kMinChunkBits);
    it.count = 1ull << (n - it.chunk_bits);
    const bool odd = (chain & kChainOdd) != 0;
    it.partial = scratch + (odd ? kPartialSpan : 0);
    it.partial_next = scratch + (odd ? 0 : kPartialSpan);
    it.variant = variant;
    it.flags = chain;
    it.fmask = fmask;
    it.fvalue = fvalue;
    return it;
}

hipError_t launch_flip_items(int dtype, const OpItem& one, const OpItem* d_items, int nitems, uint64_t total,
                             hipStream_t stream) {
    const unsigned grid = grid_for(total);
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(flip_kernel<float>, dim3(grid), dim3(256), 0, stream, one, d_items, nitems, total);
    else
        hipLaunchKernelGGL(flip_kernel<double>, dim3(grid), dim3(256), 0, stream, one, d_items, nitems, total);
    return hipGetLastError();
}

hipError_t launch_diffuse_items(int dtype, const OpItem& one, const OpItem* d_items, int nitems, uint64_t total,
                                hipStream_t stream) {
    const unsigned grid = static_cast<unsigned>(total < 65536 ? total : 65536);
    // a lone item whose sums are already there needs no pass 1
    const bool partial = d_items != nullptr || (one.flags & kChainReady) == 0;
    if (dtype == BFLY_C64) {
        if (partial)
            hipLaunchKernelGGL(diffuse_partial_kernel<float>, dim3(grid), dim3(kDiffChunkThreads), 0, stream, one,
                               d_items, nitems, total);
        hipLaunchKernelGGL(diffuse_apply_kernel<float>, dim3(grid), dim3(kDiffChunkThreads), 0, stream, one, d_items,
                           nitems, total);
    } else {
        if (partial)
            hipLaunchKernelGGL(diffuse_partial_kernel<double>, dim3(grid), dim3(kDiffChunkThreads), 0, stream, one,
                               d_items, nitems, total);
        hipLaunchKernelGGL(diffuse_apply_kernel<double>, dim3(grid), dim3(kDiffChunkThreads), 0, stream, one,
                           d_items, nitems, total);
    }
    return hipGetLastError();
}

hipError_t launch_flip(int dtype, void* re, void* im, int n, uint64_t mask, uint64_t value, hipStream_t stream) {
    const OpItem it = flip_item(re, im, n, mask, value);
    return launch_flip_items(dtype, it, nullptr, 0, it.count, stream);
}

hipError_t launch_diffuse(int dtype, void* re, void* im, int n, uint64_t mask, int variant, double* scratch,
                          int32_t chain, uint64_t fmask, uint64_t fvalue, hipStream_t stream) {
    const uint64_t size = 1ull << n;
    const uint64_t all = size - 1;
    const int t = __builtin_popcountll(mask);
    if (mask == all) {
        const OpItem it = diffuse_item(re, im, n, variant, scratch, chain, fmask, fvalue);
        return launch_diffuse_items(dtype, it, nullptr, 0, it.count, stream);
    }
    const uint64_t nblocks = 1ull << (n - t);
    if (t > kDiffuseMaxBits) {
        // wide partial mask: 2^lb blocks of 2^t amplitudes; chunks of at least
        // 2^12, at most 2^10 per block and 2^14 units (one partial buffer)
        const int lb = n - t;
        int lch = 14 - lb;
        if (lch > 10) lch = 10;
        if (lch > t - 12) lch = t - 12;
        const uint64_t others = all & ~mask;
        if (lch <= 0) {
            const unsigned grid = static_cast<unsigned>(nblocks < 65536 ? nblocks : 65536);
            if (dtype == BFLY_C64)
                hipLaunchKernelGGL(diffuse_wide_block_kernel<float>, dim3(grid), dim3(kDiffChunkThreads), 0, stream,
                                   static_cast<float*>(re), static_cast<float*>(im), t, mask, others, nblocks,
                                   variant);
            else
                hipLaunchKernelGGL(diffuse_wide_block_kernel<double>, dim3(grid), dim3(kDiffChunkThreads), 0, stream,
                                   static_cast<double*>(re), static_cast<double*>(im), t, mask, others, nblocks,
                                   variant);
            return hipGetLastError();
        }
        const int cb = t - lch;
        const uint64_t units = nblocks << lch;   // <= 2^14: 4 doubles each fit kPartialSpan
        const unsigned grid = static_cast<unsigned>(units);
        if (dtype == BFLY_C64) {
            hipLaunchKernelGGL(diffuse_wide_partial_kernel<float>, dim3(grid), dim3(kDiffChunkThreads), 0, stream,
                               static_cast<const float*>(re), static_cast<const float*>(im), cb, lch, mask, others,
                               units, scratch);
            hipLaunchKernelGGL(diffuse_wide_apply_kernel<float>, dim3(grid), dim3(kDiffChunkThreads), 0, stream,
                               static_cast<float*>(re), static_cast<float*>(im), t, cb, lch, mask, others, units,
                               static_cast<const double*>(scratch), variant);
        } else {
            hipLaunchKernelGGL(diffuse_wide_partial_kernel<double>, dim3(grid), dim3(kDiffChunkThreads), 0, stream,
                               static_cast<const double*>(re), static_cast<const double*>(im), cb, lch, mask, others,
                               units, scratch);
            hipLaunchKernelGGL(diffuse_wide_apply_kernel<double>, dim3(grid), dim3(kDiffChunkThreads), 0, stream,
                               static_cast<double*>(re), static_cast<double*>(im), t, cb, lch, mask, others, units,
                               static_cast<const double*>(scratch), variant);
        }
        return hipGetLastError();
    }
    const unsigned grid = static_cast<unsigned>(nblocks < 65536 ? nblocks : 65536);
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(diffuse_block_kernel<float>, dim3(grid), dim3(kDiffThreads), 0, stream,
                           static_cast<float*>(re), static_cast<float*>(im), t, mask, all & ~mask, nblocks, variant);
    else
        hipLaunchKernelGGL(diffuse_block_kernel<double>, dim3(grid), dim3(kDiffThreads), 0, stream,
                           static_cast<double*>(re), static_cast<double*>(im), t, mask, all & ~mask, nblocks, variant);
    return hipGetLastError();
}

hipError_t launch_marginal(int dtype, const void* re, const void* im, uint64_t size, const int* targets, int m,
                           double* scratch, hipStream_t stream) {
    if (m < 1 || m > kMargMaxBits) return hipErrorInvalidValue;
    int t[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 0; q < m; ++q) t[q] = targets[q];
    const unsigned nb = grid_for(size, kMargThreads * 16, kRedBlocks);
    double* part = scratch + (1 << kMargMaxBits);
    if (dtype == BFLY_C64)
        hipLaunchKernelGGL(marginal_kernel<float>, dim3(nb), dim3(kMargThreads), 0, stream,
                           static_cast<const float*>(re), static_cast<const float*>(im), size, m, t[0], t[1], t[2],
                           t[3], t[4], t[5], part);
    else
        hipLaunchKernelGGL(marginal_kernel<double>, dim3(nb), dim3(kMargThreads), 0, stream,
                           static_cast<const double*>(re), static_cast<const double*>(im), size, m, t[0], t[1], t[2],
                           t[3], t[4], t[5], part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kRedThreads), 0, stream, part, static_cast<int>(nb), 1 << m,
                       scratch);
    return hipGetLastError();
}

}  // namespace bfly
