// tile_ops.h -- register-group operations of the tile-pass sweeps, shared by
// the interpreting tile kernels (kernels.hip) and the straight-line pass
// kernels generated at plan time (jit.cpp, compiled with hipRTC).  No
// includes and no fixed-width typedefs: the text is handed to hipRTC as is.
//
// A thread's register group holds G = 2^S amplitudes (ar[m], ai[m]) spanned
// by the sweep's S slot qubits; member m has slot bit a equal to bit a of m.
#pragma once

namespace bfly {

__device__ __forceinline__ unsigned insert_zero32(unsigned x, int pos) {
    const unsigned lo = x & ((1u << pos) - 1);
    return ((x >> pos) << (pos + 1)) | lo;
}

// XOR swizzle of an LDS tile index (elements of elem_bytes bytes); the same
// map as lds_swizzle in bfly_plan.h.
__device__ __forceinline__ unsigned tile_swizzle(unsigned j, int elem_bytes) {
    return elem_bytes == 8 ? j ^ ((j >> 4) & 15u) : j ^ ((j >> 5) & 31u);
}

// process_pair_scalar (transformations_scalar.mojo:10-40).
template <typename T>
__device__ __forceinline__ void pair_gen(T& ur, T& ui, T& vr, T& vi, const T* g) {
    const T a = ur, b = ui, c = vr, d = vi;
    ur = a * g[0] - b * g[1] + c * g[2] - d * g[3];
    ui = a * g[1] + b * g[0] + c * g[3] + d * g[2];
    vr = a * g[4] - b * g[5] + c * g[6] - d * g[7];
    vi = a * g[5] + b * g[4] + c * g[7] + d * g[6];
}


// v *= (dr + i di) as in transform_p_scalar (transformations_scalar.mojo:330-333).
template <typename T>
__device__ __forceinline__ void cmul(T& vr, T& vi, T dr, T di) {
    const T a = vr, b = vi;
    vr = a * dr - b * di;
    vi = a * di + b * dr;
}


// A 2x2 matrix on slot SLOT of a 16-amplitude register group: pairs
// (m0, m0 | 2^SLOT).  Pairs whose slot controls do not hold take the
// identity; the test is on the member index only, so the selection is scalar
// and no register array is copied.
template <typename T, int G, int SLOT>
__device__ __forceinline__ void slot_pairs_all(T* ar, T* ai, const T* g) {
#pragma unroll
    for (int p = 0; p < G / 2; ++p) {
        constexpr int lo_mask = (1 << SLOT) - 1;
        const int m0 = ((p & ~lo_mask) << 1) | (p & lo_mask);
        pair_gen(ar[m0], ai[m0], ar[m0 | (1 << SLOT)], ai[m0 | (1 << SLOT)], g);
    }
}

template <typename T, int G, int SLOT>
__device__ __forceinline__ void slot_pairs(T* ar, T* ai, unsigned cslots, const T* g) {
    if (cslots == 0) {  // no slot controls: every pair, no selection
        slot_pairs_all<T, G, SLOT>(ar, ai, g);
        return;
    }
#pragma unroll
    for (int p = 0; p < G / 2; ++p) {
        constexpr int lo_mask = (1 << SLOT) - 1;
        const int m0 = ((p & ~lo_mask) << 1) | (p & lo_mask);
        const int m1 = m0 | (1 << SLOT);
        const bool on = (static_cast<unsigned>(m0) & cslots) == cslots;
        T h[8];
        h[0] = on ? g[0] : T(1);
        h[1] = on ? g[1] : T(0);
        h[2] = on ? g[2] : T(0);
        h[3] = on ? g[3] : T(0);
        h[4] = on ? g[4] : T(0);
        h[5] = on ? g[5] : T(0);
        h[6] = on ? g[6] : T(1);
        h[7] = on ? g[7] : T(0);
        pair_gen(ar[m0], ai[m0], ar[m1], ai[m1], h);
    }
}

// Fused multiply-add in the state's precision (a single rounding).
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Z family: multiply the members selected in sel1 by -1 and the others by 1
// (exact: the products flip or keep the sign bit).  sel1 is the same in
// every lane, so the factors are scalars; whole-register products keep the
// register arrays in place, where a sign-bit XOR on the high half of each
// double costs a copy of the low half.
template <typename T, int G>
__device__ __forceinline__ void neg_apply(T* ar, T* ai, unsigned sel1) {
#pragma unroll
    for (int m = 0; m < G; ++m) {
        const T f = ((sel1 >> m) & 1u) ? T(-1) : T(1);
        ar[m] *= f;
        ai[m] *= f;
    }
}

// Z family with the target bit outside the slots: `bit` (per lane or per
// tile) is the target's value; members in mm1 flip when it is 1, members in
// mm0 when it is 0 (mm1 and mm0 are the same in every lane).  With
// F = -1 where bit is 1 and 1 elsewhere, member m's factor is a F + b with
// scalar (a, b): (1, 0) for mm1 only, (-1, 0) for mm0 only, (0, -1) for
// both, (0, 1) for neither; every value is exact.
template <typename T, int G>
__device__ __forceinline__ void neg_apply_bit(T* ar, T* ai, unsigned mm1, unsigned mm0, unsigned bit) {
    const T F = bit ? T(-1) : T(1);
#pragma unroll
    for (int m = 0; m < G; ++m) {
        const unsigned f1 = (mm1 >> m) & 1u, f0 = (mm0 >> m) & 1u;
        const T a = f1 == f0 ? T(0) : (f1 ? T(1) : T(-1));
        const T b = f1 != f0 ? T(0) : (f1 ? T(-1) : T(1));
        const T f = fmat(a, F, b);
        ar[m] *= f;
        ai[m] *= f;
    }
}

// Phase family: members in sel1 are multiplied by d1, in sel0 by d0, the
// others keep their value (factor 1 + 0i, exact).
template <typename T, int G>
__device__ __forceinline__ void diag_apply(T* ar, T* ai, unsigned sel1, unsigned sel0, const T* g) {
#pragma unroll
    for (int m = 0; m < G; ++m) {
        const bool s1 = (sel1 >> m) & 1u;
        const bool s0 = (sel0 >> m) & 1u;
        const T fr = s1 ? g[2] : (s0 ? g[0] : T(1));
        const T fi = s1 ? g[3] : (s0 ? g[1] : T(0));
        cmul(ar[m], ai[m], fr, fi);
    }
}

__device__ __forceinline__ double fabs_t(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float fabs_t(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ bool signbit_t(double x) { return __builtin_signbit(x); }
__device__ __forceinline__ bool signbit_t(float x) { return __builtin_signbitf(x); }



// T_ROT on slot SLOT: v <- v * (zr + i zi) (the pre-phase, exact when it
// is 1 or -1), then the scaled real rotation of a factored one-qubit
// unitary, one FMA per output component: form 0: (u, v) <- (u - t v, v + t u);
// form 1: (u, v) <- (t u - v, u + t v).
// One pair of a T_ROT: v <- v * (zr + i zi), then form 0:
// (u, v) <- (u - t v, v + t u), or form 1: (u, v) <- (t u - v, u + t v).
template <typename T>
__device__ __forceinline__ void rot_pair0(T& ur_, T& ui_, T& vr_, T& vi_, T t, T zr, T zi) {
    const T ur = ur_, ui = ui_;
    const T vr = vr_ * zr - vi_ * zi, vi = vr_ * zi + vi_ * zr;
    const T nt = -t;
    ur_ = fmat(nt, vr, ur);
    ui_ = fmat(nt, vi, ui);
    vr_ = fmat(t, ur, vr);
    vi_ = fmat(t, ui, vi);
}
template <typename T>
__device__ __forceinline__ void rot_pair1(T& ur_, T& ui_, T& vr_, T& vi_, T t, T zr, T zi) {
    const T ur = ur_, ui = ui_;
    const T vr = vr_ * zr - vi_ * zi, vi = vr_ * zi + vi_ * zr;
    ur_ = fmat(t, ur, -vr);
    ui_ = fmat(t, ui, -vi);
    vr_ = fmat(t, vr, ur);
    vi_ = fmat(t, vi, ui);
}

// x with its sign bit XORed with bit 31 of s (s = 0 or 0x80000000u): a
// per-lane sign flip in one integer instruction.
__device__ __forceinline__ double xsign(double x, unsigned s) {
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) ^ (static_cast<unsigned long long>(s) << 32));
}
__device__ __forceinline__ float xsign(float x, unsigned s) {
    return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, x) ^ s);
}

// The same rotations with the pre-phase z = SZ (+1 or -1, a compile-time
// sign): the generated kernels apply a sweep's pre-phases earlier, as one
// factor per register member, and a post flip of v leaves z = -1.  Four FMAs
// per pair; the sign folds into the FMA operands.
template <int SZ, typename T>
__device__ __forceinline__ void rot_pair0z(T& ur_, T& ui_, T& vr_, T& vi_, T t) {
    const T ur = ur_, ui = ui_;
    const T vr = SZ > 0 ? vr_ : -vr_, vi = SZ > 0 ? vi_ : -vi_;
    const T nt = -t;
    ur_ = fmat(nt, vr, ur);
    ui_ = fmat(nt, vi, ui);
    vr_ = fmat(t, ur, vr);
    vi_ = fmat(t, ui, vi);
}
template <int SZ, typename T>
__device__ __forceinline__ void rot_pair1z(T& ur_, T& ui_, T& vr_, T& vi_, T t) {
    const T ur = ur_, ui = ui_;
    const T vr = SZ > 0 ? vr_ : -vr_, vi = SZ > 0 ? vi_ : -vi_;
    ur_ = fmat(t, ur, -vr);
    ui_ = fmat(t, ui, -vi);
    vr_ = fmat(t, vr, ur);
    vi_ = fmat(t, vi, ui);
}

template <typename T, int G, int SLOT>
__device__ __forceinline__ void slot_rot(T* ar, T* ai, T t, bool form1, T zr, T zi) {
    constexpr int lo_mask = (1 << SLOT) - 1;
    if (!form1) {
#pragma unroll
        for (int p = 0; p < G / 2; ++p) {
            const int m0 = ((p & ~lo_mask) << 1) | (p & lo_mask);
            const int m1 = m0 | (1 << SLOT);
            rot_pair0(ar[m0], ai[m0], ar[m1], ai[m1], t, zr, zi);
        }
    } else {
#pragma unroll
        for (int p = 0; p < G / 2; ++p) {
            const int m0 = ((p & ~lo_mask) << 1) | (p & lo_mask);
            const int m1 = m0 | (1 << SLOT);
            rot_pair1(ar[m0], ai[m0], ar[m1], ai[m1], t, zr, zi);
        }
    }
}

// T_PH on slot SLOT: members whose slot bit is 1 times (zr + i zi).
template <typename T, int G, int SLOT>
__device__ __forceinline__ void slot_phase(T* ar, T* ai, T zr, T zi) {
#pragma unroll
    for (int m = 0; m < G; ++m)
        if ((m >> SLOT) & 1) cmul(ar[m], ai[m], zr, zi);
}

// T_SCALE: every member times (zr + i zi); a real factor takes two products.
template <typename T, int G>
__device__ __forceinline__ void scale_all(T* ar, T* ai, T zr, T zi) {
    if (zi == T(0)) {
#pragma unroll
        for (int m = 0; m < G; ++m) {
            ar[m] *= zr;
            ai[m] *= zr;
        }
    } else {
#pragma unroll
        for (int m = 0; m < G; ++m) cmul(ar[m], ai[m], zr, zi);
    }
}


}  // namespace bfly
