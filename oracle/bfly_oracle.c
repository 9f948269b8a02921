/*
 * bfly_oracle.c -- CPU restatement of the Butterfly (learnqc/code_mojo)
 * SCALAR execution strategy, used ONLY as test infrastructure.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the
 * timed CPU baseline.  The product (libbfly_hip.so, code_mojo_amd/) never
 * links, loads or calls it.
 *
 * It consumes the same lowered op array (bfly_op, include/bfly_hip.h) as the
 * GPU strategy and follows the reference's scalar kernels step by step:
 *   dispatch            butterfly/core/executors.mojo:268-319, 331-372
 *   validation          butterfly/core/executors.mojo:90-185
 *   1-qubit kernels     butterfly/core/transformations_scalar.mojo:10-40, 64-117,
 *                       178-394 (X, H, P, RY specialisations, generic pair)
 *   controlled kernels  butterfly/core/transformations_scalar.mojo:121-175, 397-566
 *   multi-control       butterfly/core/transformations_scalar.mojo:569-594
 *   dense unitaries     butterfly/core/transformations_unitary.mojo:22-139
 *   bit reversal        butterfly/core/state.mojo:131-204 (full), 252-301 (partial)
 *   SWAP via 3 CX       butterfly/core/executors.mojo:231-241
 *   phase flip, diffuser butterfly/algos/grover_classical.mojo:262-326
 *   measurement         butterfly/core/state.mojo:447-553
 * Build with -O2 -ffp-contract=off so that no FMA contraction changes the
 * rounding of the reference's separate multiply/add expressions.
 *
 * Compiled twice: FLOAT_T=double (reference default) and FLOAT_T=float (the
 * reference's `-D FLOAT32` build, butterfly/core/types.mojo:7).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/bfly_hip.h"

#ifndef FLOAT_T
#define FLOAT_T double
#define SFX(name) name##_f64
#endif

typedef FLOAT_T real;

static void set_err(char* err, int errlen, const char* msg) {
    if (err && errlen > 0) {
        strncpy(err, msg, (size_t)errlen - 1);
        err[errlen - 1] = 0;
    }
}

/* process_pair_scalar: generic 2x2 complex mat-vec on (k0, k1)
   (transformations_scalar.mojo:10-40). */
static inline void SFX(pair_generic)(real* re, real* im, const real g[8], uint64_t k0, uint64_t k1) {
    real u_re = re[k0], u_im = im[k0], v_re = re[k1], v_im = im[k1];
    re[k0] = u_re * g[0] - u_im * g[1] + v_re * g[2] - v_im * g[3];
    im[k0] = u_re * g[1] + u_im * g[0] + v_re * g[3] + v_im * g[2];
    re[k1] = u_re * g[4] - u_im * g[5] + v_re * g[6] - v_im * g[7];
    im[k1] = u_re * g[5] + u_im * g[4] + v_re * g[7] + v_im * g[6];
}

/* Pair walk shared by every kernel: block b of size 2*stride, offset m
   (the `idx = 2*j - r` walk of transformations_scalar.mojo:111-117 visits the
   same pairs in the same order). */
#define FOR_PAIRS(L, stride, idx)                                   \
    for (uint64_t _b = 0; _b < (L); _b += 2 * (stride))             \
        for (uint64_t _m = 0, idx = _b; _m < (stride); ++_m, idx = _b + _m)

/* transform_x_scalar (transformations_scalar.mojo:178-230) with an optional
   single-control test (c_transform_x_scalar, :397-452). */
static void SFX(kernel_x)(real* re, real* im, uint64_t L, uint64_t stride, int64_t control) {
    FOR_PAIRS(L, stride, idx) {
        if (control >= 0 && !((idx >> control) & 1)) continue;
        real u_re = re[idx], u_im = im[idx];
        re[idx] = re[idx + stride];
        im[idx] = im[idx + stride];
        re[idx + stride] = u_re;
        im[idx + stride] = u_im;
    }
}

/* transform_h_scalar (transformations_scalar.mojo:233-286): the constant is
   the literal 0.7071067811865476 (:241). */
static void SFX(kernel_h)(real* re, real* im, uint64_t L, uint64_t stride) {
    const real sq_half = (real)0.7071067811865476;
    FOR_PAIRS(L, stride, idx) {
        real u_re = re[idx], u_im = im[idx], v_re = re[idx + stride], v_im = im[idx + stride];
        re[idx] = (u_re + v_re) * sq_half;
        im[idx] = (u_im + v_im) * sq_half;
        re[idx + stride] = (u_re - v_re) * sq_half;
        im[idx + stride] = (u_im - v_im) * sq_half;
    }
}

/* transform_p_scalar (:289-336) and c_transform_p_scalar (:455-505). */
static void SFX(kernel_p)(real* re, real* im, uint64_t L, uint64_t stride, real theta, int64_t control) {
    real cos_t = (real)cos((double)theta), sin_t = (real)sin((double)theta);
    FOR_PAIRS(L, stride, idx) {
        if (control >= 0 && !((idx >> control) & 1)) continue;
        uint64_t k = idx + stride;
        real v_re = re[k], v_im = im[k];
        re[k] = v_re * cos_t - v_im * sin_t;
        im[k] = v_re * sin_t + v_im * cos_t;
    }
}

/* transform_ry_scalar (:339-394) and c_transform_ry_scalar (:508-566). */
static void SFX(kernel_ry)(real* re, real* im, uint64_t L, uint64_t stride, real theta, int64_t control) {
    real cos_t = (real)cos((double)(theta / 2)), sin_t = (real)sin((double)(theta / 2));
    FOR_PAIRS(L, stride, idx) {
        if (control >= 0 && !((idx >> control) & 1)) continue;
        real u_re = re[idx], u_im = im[idx], v_re = re[idx + stride], v_im = im[idx + stride];
        re[idx] = u_re * cos_t - v_re * sin_t;
        im[idx] = u_im * cos_t - v_im * sin_t;
        re[idx + stride] = u_re * sin_t + v_re * cos_t;
        im[idx + stride] = u_im * sin_t + v_im * cos_t;
    }
}

/* transform_scalar generic loop (:111-117), c_transform_scalar (:167-175) and
   mc_transform_scalar (:586-594): process_pair_scalar where
   (idx & mask) == mask. */
static void SFX(kernel_generic)(real* re, real* im, uint64_t L, uint64_t stride, const real g[8], uint64_t mask) {
    FOR_PAIRS(L, stride, idx) {
        if ((idx & mask) != mask) continue;
        SFX(pair_generic)(re, im, g, idx, idx + stride);
    }
}

static int popcount64(uint64_t x) {
    int c = 0;
    while (x) { x &= x - 1; ++c; }
    return c;
}

static int lowest_bit(uint64_t x) {
    int b = 0;
    while (!((x >> b) & 1)) ++b;
    return b;
}

/* handle_gate_scalar (executors.mojo:331-366) with the default ExecContext:
   every simd_use_specialized_* flag is True (utils/context.mojo:4-10,105-111). */
static void SFX(apply_gate)(int n, real* re, real* im, const bfly_op* op) {
    uint64_t L = 1ull << n;
    uint64_t stride = 1ull << op->target;
    real g[8];
    for (int i = 0; i < 8; ++i) g[i] = (real)op->mat[i];
    int nctrl = popcount64(op->controls);
    real theta = (real)op->theta;
    if (nctrl == 0) {
        /* transform_scalar (transformations_scalar.mojo:64-117) */
        switch (op->gate) {
            case BFLY_GATE_H: SFX(kernel_h)(re, im, L, stride); return;
            case BFLY_GATE_P: SFX(kernel_p)(re, im, L, stride, theta, -1); return;
            case BFLY_GATE_X: SFX(kernel_x)(re, im, L, stride, -1); return;
            case BFLY_GATE_RY: SFX(kernel_ry)(re, im, L, stride, theta, -1); return;
            default: SFX(kernel_generic)(re, im, L, stride, g, 0); return;
        }
    } else if (nctrl == 1) {
        /* c_transform_scalar (transformations_scalar.mojo:121-175) */
        int c = lowest_bit(op->controls);
        switch (op->gate) {
            case BFLY_GATE_P: SFX(kernel_p)(re, im, L, stride, theta, c); return;
            case BFLY_GATE_X: SFX(kernel_x)(re, im, L, stride, c); return;
            case BFLY_GATE_RY: SFX(kernel_ry)(re, im, L, stride, theta, c); return;
            default: SFX(kernel_generic)(re, im, L, stride, g, op->controls); return;
        }
    } else {
        /* mc_transform_scalar (transformations_scalar.mojo:569-594): generic
           matrix for every gate kind. */
        SFX(kernel_generic)(re, im, L, stride, g, op->controls);
    }
}

/* apply_unitary / apply_controlled_unitary (transformations_unitary.mojo:22-139),
   with the single control generalised to a control mask (the reference's
   multi-control c_unitary drops the matrix, circuit.mojo:786-808; that defect
   is not reproduced). */
static int SFX(apply_unitary)(int n, real* re, real* im, const double* u, int target, int m, uint64_t ctrl) {
    uint64_t dim = 1ull << m;
    uint64_t size = 1ull << n;
    uint64_t stride = 1ull << target;
    uint64_t block = 1ull << (target + m);
    real* tmp_re = (real*)malloc(sizeof(real) * dim * 4);
    if (!tmp_re) return BFLY_ENOMEM;
    real* tmp_im = tmp_re + dim;
    real* out_re = tmp_im + dim;
    real* out_im = out_re + dim;
    for (uint64_t high = 0; high < size; high += block) {
        for (uint64_t low = 0; low < stride; ++low) {
            uint64_t base = high + low;
            if ((base & ctrl) != ctrl) continue;
            for (uint64_t k = 0; k < dim; ++k) {
                uint64_t idx = base + (k << target);
                tmp_re[k] = re[idx];
                tmp_im[k] = im[idx];
            }
            for (uint64_t row = 0; row < dim; ++row) {
                real sum_re = 0, sum_im = 0;
                for (uint64_t col = 0; col < dim; ++col) {
                    real ur = (real)u[2 * (row * dim + col)], ui = (real)u[2 * (row * dim + col) + 1];
                    real v_re = tmp_re[col], v_im = tmp_im[col];
                    sum_re += ur * v_re - ui * v_im;
                    sum_im += ur * v_im + ui * v_re;
                }
                out_re[row] = sum_re;
                out_im[row] = sum_im;
            }
            for (uint64_t k = 0; k < dim; ++k) {
                uint64_t idx = base + (k << target);
                re[idx] = out_re[k];
                im[idx] = out_im[k];
            }
        }
    }
    free(tmp_re);
    return BFLY_OK;
}

static uint64_t bit_reverse_n(uint64_t i, int nbits) {
    uint64_t r = 0;
    for (int b = 0; b < nbits; ++b) r |= ((i >> b) & 1ull) << (nbits - 1 - b);
    return r;
}

/* bit_reverse_state (state.mojo:131-204): out of place, out[i] = in[rev(i)],
   then the lists are swapped in. */
static int SFX(bitrev_full)(int n, real* re, real* im) {
    uint64_t size = 1ull << n;
    real* s_re = (real*)malloc(sizeof(real) * size);
    real* s_im = (real*)malloc(sizeof(real) * size);
    if (!s_re || !s_im) { free(s_re); free(s_im); return BFLY_ENOMEM; }
    for (uint64_t i = 0; i < size; ++i) {
        uint64_t r = bit_reverse_n(i, n);
        s_re[i] = re[r];
        s_im[i] = im[r];
    }
    memcpy(re, s_re, sizeof(real) * size);
    memcpy(im, s_im, sizeof(real) * size);
    free(s_re);
    free(s_im);
    return BFLY_OK;
}

/* partial_bit_reverse_state (state.mojo:252-301): targets sorted ascending,
   swap i <-> j where the target-bit sub-index is reversed and i_small < j_small. */
static void SFX(bitrev_partial)(int n, real* re, real* im, uint64_t mask) {
    int t[64], k = 0;
    for (int b = 0; b < n; ++b) if ((mask >> b) & 1) t[k++] = b;
    if (k <= 1) return;
    uint64_t size = 1ull << n;
    for (uint64_t i_large = 0; i_large < size; ++i_large) {
        uint64_t i_small = 0;
        for (int b = 0; b < k; ++b) if ((i_large >> t[b]) & 1) i_small |= 1ull << b;
        uint64_t j_small = 0;
        for (int b = 0; b < k; ++b) if ((i_small >> b) & 1) j_small |= 1ull << (k - 1 - b);
        if (i_small < j_small) {
            uint64_t j_large = i_large;
            for (int b = 0; b < k; ++b) {
                if ((j_small >> b) & 1) j_large |= 1ull << t[b];
                else j_large &= ~(1ull << t[b]);
            }
            real tr = re[i_large], ti = im[i_large];
            re[i_large] = re[j_large];
            im[i_large] = im[j_large];
            re[j_large] = tr;
            im[j_large] = ti;
        }
    }
}

/* apply_bit_reverse (state.mojo:383-390): empty target list -> full reversal,
   otherwise the partial one.  The lowered op carries the targets as a mask;
   a mask equal to all n bits takes the partial route exactly as a full target
   list does in the reference (the two agree: tests/test_bit_reverse.mojo:23-34). */
static int SFX(apply_bitrev)(int n, real* re, real* im, uint64_t mask) {
    if (mask == 0) return SFX(bitrev_full)(n, re, im);
    SFX(bitrev_partial)(n, re, im, mask);
    return BFLY_OK;
}

/* apply_swap_with_cx (executors.mojo:231-241) with handle_cx_scalar
   (executors.mojo:369-372): three controlled X. */
static void SFX(apply_swap)(int n, real* re, real* im, int a, int b) {
    if (a == b) return;
    uint64_t L = 1ull << n;
    SFX(kernel_x)(re, im, L, 1ull << b, a);
    SFX(kernel_x)(re, im, L, 1ull << a, b);
    SFX(kernel_x)(re, im, L, 1ull << b, a);
}

/* Scatter the bits of x into the set positions of mask (ascending):
   _base_index_from_bits (algos/grover_classical.mojo:248-259). */
static uint64_t deposit_bits(uint64_t x, uint64_t mask) {
    uint64_t r = 0;
    for (int b = 0; b < 64 && mask; ++b) {
        if (!((mask >> b) & 1)) continue;
        if (x & 1) r |= 1ull << b;
        x >>= 1;
        mask &= ~(1ull << b);
    }
    return r;
}

/* PHASE_FLIP: classical_oracle (algos/grover.mojo:64-66) negates listed
   indices; apply_grover_inversion_0 (algos/grover_classical.mojo:262-281)
   negates the target-bits-zero amplitude of every block.  Both are "negate
   where (i & mask) == value". */
static void SFX(apply_flip)(int n, real* re, real* im, uint64_t mask, uint64_t value) {
    uint64_t all = (1ull << n) - 1, others = all & ~mask;
    uint64_t blocks = 1ull << (n - popcount64(mask));
    for (uint64_t b = 0; b < blocks; ++b) {
        uint64_t idx = deposit_bits(b, others) | value;
        re[idx] = -re[idx];
        im[idx] = -im[idx];
    }
}

/* DIFFUSE: apply_grover_diffuser (algos/grover_classical.mojo:284-326):
   per block, naive sums over the target subspace, add = 2/2^t * sum;
   variant 1 keeps the reference's update (re <- add_re - re,
   im <- -2/2^t * sum_im - im); variant 0 is the action of the gate
   diffuser H X MCP(pi) X H (algos/grover.mojo:45-62): a <- a - add. */
static void SFX(apply_diffuse)(int n, real* re, real* im, uint64_t mask, int variant) {
    uint64_t all = (1ull << n) - 1, others = all & ~mask;
    int t = popcount64(mask);
    uint64_t block_size = 1ull << t, blocks = 1ull << (n - t);
    real inv_block = (real)1 / (real)block_size;
    for (uint64_t b = 0; b < blocks; ++b) {
        uint64_t base = deposit_bits(b, others);
        real sum_re = 0, sum_im = 0;
        for (uint64_t e = 0; e < block_size; ++e) {
            uint64_t idx = base | deposit_bits(e, mask);
            sum_re += re[idx];
            sum_im += im[idx];
        }
        real add_re = (real)2 * inv_block * sum_re;
        real add_im = (variant == 1 ? (real)-2 : (real)2) * inv_block * sum_im;
        for (uint64_t e = 0; e < block_size; ++e) {
            uint64_t idx = base | deposit_bits(e, mask);
            if (variant == 1) {
                re[idx] = add_re - re[idx];
                im[idx] = add_im - im[idx];
            } else {
                re[idx] = re[idx] - add_re;
                im[idx] = im[idx] - add_im;
            }
        }
    }
}

/* MEASURE: apply_measure (state.mojo:447-553) step by step: outcome
   probabilities of the free targets among the amplitudes matching the forced
   values, summed naively in index order in the state's precision; one MINSTD
   step from the seed; the first outcome whose running sum reaches r; the
   matching amplitudes times 1 / sqrt(p), the others 0.  `tv` pairs are the
   op's udata entries: (qubit, value), value -1 = free. */
static int SFX(apply_measure)(int n, real* re, real* im, const double* tv, int m, int64_t seed, char* err,
                              int errlen) {
    if (m == 0) return BFLY_OK;
    int targets[64], values[64], free_targets[64], free_pos[64];
    int nfree = 0;
    if (m > 64) { set_err(err, errlen, "too many measurement targets"); return BFLY_EINVAL; }
    for (int j = 0; j < m; ++j) {
        targets[j] = (int)tv[2 * j];
        values[j] = (int)tv[2 * j + 1];
        if (targets[j] < 0 || targets[j] >= n) { set_err(err, errlen, "Target out of bounds"); return BFLY_EINVAL; }
    }
    for (int j = 0; j < m; ++j) {
        if (values[j] >= 0) {
            free_pos[j] = -1;
        } else {
            free_pos[j] = nfree;
            free_targets[nfree++] = targets[j];
        }
    }
    /* 2^nfree outcome sums in host memory, as the reference allocates them */
    if (nfree > 30) { set_err(err, errlen, "too many free measurement targets"); return BFLY_EUNSUPPORTED; }
    uint64_t size = 1ull << n, outcome_count = 1ull << nfree;
    real* probs = (real*)calloc(outcome_count, sizeof(real));
    if (!probs) { set_err(err, errlen, "allocation failed"); return BFLY_ENOMEM; }
    real total = 0;
    for (uint64_t i = 0; i < size; ++i) {
        int is_match = 1;
        for (int j = 0; j < m; ++j) {
            if (values[j] >= 0 && (int)((i >> targets[j]) & 1) != values[j]) { is_match = 0; break; }
        }
        if (!is_match) continue;
        uint64_t outcome = 0;
        for (int k = 0; k < nfree; ++k)
            if ((i >> free_targets[k]) & 1) outcome |= 1ull << k;
        real p = re[i] * re[i] + im[i] * im[i];
        probs[outcome] = probs[outcome] + p;
        total += p;
    }
    if (total <= 0) { free(probs); set_err(err, errlen, "Measurement probability is zero"); return BFLY_EINVAL; }
    int64_t rng = seed % 2147483647;
    rng = (rng * 48271) % 2147483647;
    real r = (real)rng / (real)2147483647 * total;
    uint64_t chosen = 0;
    real acc = 0;
    for (uint64_t o = 0; o < outcome_count; ++o) {
        acc += probs[o];
        if (r <= acc) { chosen = o; break; }
    }
    real chosen_prob = probs[chosen];
    free(probs);
    if (chosen_prob <= 0) { set_err(err, errlen, "Measurement outcome probability is zero"); return BFLY_EINVAL; }
    real inv = (real)1.0 / (real)sqrt((double)chosen_prob);
    for (uint64_t i = 0; i < size; ++i) {
        int is_match = 1;
        for (int j = 0; j < m; ++j) {
            int desired = values[j] >= 0 ? values[j] : (int)((chosen >> free_pos[j]) & 1);
            if ((int)((i >> targets[j]) & 1) != desired) { is_match = 0; break; }
        }
        if (is_match) {
            re[i] = re[i] * inv;
            im[i] = im[i] * inv;
        } else {
            re[i] = 0;
            im[i] = 0;
        }
    }
    return BFLY_OK;
}

/* validate_circuit (executors.mojo:129-185) for the lowered op kinds. */
static int validate_ops(int n, const bfly_op* ops, uint64_t nops, uint64_t udata_len, char* err, int errlen) {
    char buf[160];
    uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    for (uint64_t i = 0; i < nops; ++i) {
        const bfly_op* op = &ops[i];
        switch (op->kind) {
            case BFLY_OP_GATE:
                if (op->target < 0 || op->target >= n) {
                    snprintf(buf, sizeof buf, "Target out of bounds: %d", op->target);
                    set_err(err, errlen, buf); return BFLY_EINVAL;
                }
                if (op->controls & ~all) { set_err(err, errlen, "Control out of bounds"); return BFLY_EINVAL; }
                if ((op->controls >> op->target) & 1) { set_err(err, errlen, "Control overlaps target"); return BFLY_EINVAL; }
                break;
            case BFLY_OP_BITREV:
                if (op->aux & ~all) { set_err(err, errlen, "Target out of bounds"); return BFLY_EINVAL; }
                break;
            case BFLY_OP_SWAP:
                if (op->target < 0 || op->target >= n || op->arg_i < 0 || op->arg_i >= n) {
                    set_err(err, errlen, "Target out of bounds"); return BFLY_EINVAL;
                }
                break;
            case BFLY_OP_UNITARY: {
                int m = op->arg_i;
                if (m <= 0) break;
                if (op->target < 0 || op->target + m > n) { set_err(err, errlen, "Target range out of bounds"); return BFLY_EINVAL; }
                uint64_t range = ((1ull << m) - 1) << op->target;
                if (op->controls & ~all) { set_err(err, errlen, "Control out of bounds"); return BFLY_EINVAL; }
                if (op->controls & range) { set_err(err, errlen, "Control overlaps target range"); return BFLY_EINVAL; }
                uint64_t need = op->aux + (1ull << (2 * m));
                if (need > udata_len) { set_err(err, errlen, "Unitary size mismatch"); return BFLY_EINVAL; }
                break;
            }
            case BFLY_OP_PHASE_FLIP:
            case BFLY_OP_DIFFUSE:
                if (op->aux == 0 || (op->aux & ~all)) { set_err(err, errlen, "Target out of bounds"); return BFLY_EINVAL; }
                if (op->kind == BFLY_OP_PHASE_FLIP && (op->controls & ~op->aux)) {
                    set_err(err, errlen, "phase-flip pattern outside its mask"); return BFLY_EINVAL;
                }
                if (op->kind == BFLY_OP_DIFFUSE && op->arg_i != 0 && op->arg_i != 1) {
                    set_err(err, errlen, "unknown diffuser variant"); return BFLY_EINVAL;
                }
                break;
            case BFLY_OP_MEASURE:
                if (op->arg_i < 0 || op->aux + (uint64_t)op->arg_i > udata_len) {
                    set_err(err, errlen, "Measurement values size mismatch"); return BFLY_EINVAL;
                }
                if (op->gate != 1) { set_err(err, errlen, "the oracle needs a measurement seed"); return BFLY_EINVAL; }
                break;
            default:
                set_err(err, errlen, "Unknown transformation type");
                return BFLY_EUNSUPPORTED;
        }
    }
    return BFLY_OK;
}

/* execute_scalar (executors.mojo:585-598) over a lowered op array.
   udata holds complex unitary entries (re, im interleaved); udata_len counts
   complex entries. */
int SFX(oracle_execute)(int n, real* re, real* im, const bfly_op* ops, uint64_t nops,
                        const double* udata, uint64_t udata_len, char* err, int errlen) {
    if (n < 0 || n > 40) { set_err(err, errlen, "bad qubit count"); return BFLY_EINVAL; }
    int rc = validate_ops(n, ops, nops, udata_len, err, errlen);
    if (rc) return rc;
    for (uint64_t i = 0; i < nops; ++i) {
        const bfly_op* op = &ops[i];
        switch (op->kind) {
            case BFLY_OP_GATE: SFX(apply_gate)(n, re, im, op); break;
            case BFLY_OP_BITREV:
                rc = SFX(apply_bitrev)(n, re, im, op->aux);
                if (rc) { set_err(err, errlen, "bit reversal allocation failed"); return rc; }
                break;
            case BFLY_OP_SWAP: SFX(apply_swap)(n, re, im, op->target, op->arg_i); break;
            case BFLY_OP_UNITARY:
                if (op->arg_i <= 0) break;
                rc = SFX(apply_unitary)(n, re, im, udata + 2 * op->aux, op->target, op->arg_i, op->controls);
                if (rc) { set_err(err, errlen, "unitary allocation failed"); return rc; }
                break;
            case BFLY_OP_PHASE_FLIP: SFX(apply_flip)(n, re, im, op->aux, op->controls); break;
            case BFLY_OP_DIFFUSE: SFX(apply_diffuse)(n, re, im, op->aux, op->arg_i); break;
            case BFLY_OP_MEASURE:
                rc = SFX(apply_measure)(n, re, im, udata + 2 * op->aux, op->arg_i, (int64_t)op->controls, err, errlen);
                if (rc) return rc;
                break;
        }
    }
    return BFLY_OK;
}

/* Stand-alone entry points for the permutation tests. */
int SFX(oracle_bitrev)(int n, real* re, real* im, uint64_t mask) { return SFX(apply_bitrev)(n, re, im, mask); }
void SFX(oracle_swap_bits)(int n, real* re, real* im, int a, int b) { SFX(apply_swap)(n, re, im, a, b); }
