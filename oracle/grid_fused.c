/*
 * grid_fused.c -- OpenMP restatement of the reference's fastest CPU strategy,
 * GRID_PARALLEL_FUSED, used ONLY as the timed CPU baseline beside the GPU.
 *
 * TEST INFRASTRUCTURE: loaded only by bench.py's cpu_baseline leg and tests.
 *
 * Semantics followed (the grouping pass itself, transformations_grid_fusion.mojo,
 * is absent from the reference snapshot; its call site fixes the contract):
 *   layout        col_bits = max(n - slack, min_bits), rows = 2^(n - col_bits)
 *                 butterfly/core/executors.mojo:213-228
 *   grouping      consecutive gates whose target and controls are all
 *                 < col_bits form one row-local group applied row by row in
 *                 one pass per row; everything else runs through execute_grid
 *                 butterfly/core/executors.mojo:949-955,1062-1093
 *   row kernels   transform_row_{h,p,}_simd     transformations_grid.mojo:120-364
 *   cross-row     transform_column_{h_,}simd_tiled over 256-column tiles
 *                 transformations_grid.mojo:21,471-555,916-978
 *   controlled    cross-row / multi-controlled gates: flat parallel passes
 *                 (c_transform_p_simd_parallel, mc_transform_simd_parallel)
 *                 transformations_grid.mojo:1009-1139, executors.mojo:522-559
 *   bit reversal  parallel out-of-place gather        state.mojo:131-204
 * Gate arithmetic is the scalar strategy's (same expressions as bfly_oracle.c).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "../include/bfly_hip.h"

#define L2_TILE_COLS 256

typedef struct {
    int kind;            /* 0 = H, 1 = P, 2 = X, 3 = RY, 4 = generic */
    uint64_t stride;
    uint64_t cmask;
    double g[8];
    double c, s;         /* P: cos/sin(theta); RY: cos/sin(theta/2) */
} gate_k;

static void make_gate(const bfly_op* op, gate_k* k) {
    k->stride = 1ull << op->target;
    k->cmask = op->controls;
    memcpy(k->g, op->mat, sizeof k->g);
    int nctrl = __builtin_popcountll(op->controls);
    k->kind = 4;
    if (nctrl <= 1) {
        if (op->gate == BFLY_GATE_H && nctrl == 0) k->kind = 0;
        else if (op->gate == BFLY_GATE_P) { k->kind = 1; k->c = cos(op->theta); k->s = sin(op->theta); }
        else if (op->gate == BFLY_GATE_X) k->kind = 2;
        else if (op->gate == BFLY_GATE_RY) { k->kind = 3; k->c = cos(op->theta / 2); k->s = sin(op->theta / 2); }
    }
}

static inline void apply_pair(const gate_k* k, double* re, double* im, uint64_t i0) {
    uint64_t i1 = i0 + k->stride;
    double u_re = re[i0], u_im = im[i0], v_re = re[i1], v_im = im[i1];
    const double sq_half = 0.7071067811865476;
    switch (k->kind) {
        case 0:
            re[i0] = (u_re + v_re) * sq_half; im[i0] = (u_im + v_im) * sq_half;
            re[i1] = (u_re - v_re) * sq_half; im[i1] = (u_im - v_im) * sq_half;
            break;
        case 1:
            re[i1] = v_re * k->c - v_im * k->s; im[i1] = v_re * k->s + v_im * k->c;
            break;
        case 2:
            re[i0] = v_re; im[i0] = v_im; re[i1] = u_re; im[i1] = u_im;
            break;
        case 3:
            re[i0] = u_re * k->c - v_re * k->s; im[i0] = u_im * k->c - v_im * k->s;
            re[i1] = u_re * k->s + v_re * k->c; im[i1] = u_im * k->s + v_im * k->c;
            break;
        default: {
            const double* g = k->g;
            re[i0] = u_re * g[0] - u_im * g[1] + v_re * g[2] - v_im * g[3];
            im[i0] = u_re * g[1] + u_im * g[0] + v_re * g[3] + v_im * g[2];
            re[i1] = u_re * g[4] - u_im * g[5] + v_re * g[6] - v_im * g[7];
            im[i1] = u_re * g[5] + u_im * g[4] + v_re * g[7] + v_im * g[6];
        }
    }
}

/* A row-local group: every gate applied to one row before the next row. */
static void run_row_group(double* re, double* im, const gate_k* gs, int ng, uint64_t rows, uint64_t row_size) {
#pragma omp parallel for schedule(static)
    for (uint64_t row = 0; row < rows; ++row) {
        uint64_t off = row * row_size;
        for (int gi = 0; gi < ng; ++gi) {
            const gate_k* k = &gs[gi];
            for (uint64_t b = 0; b < row_size; b += 2 * k->stride)
                for (uint64_t m = 0; m < k->stride; ++m) {
                    uint64_t i0 = off + b + m;
                    if ((i0 & k->cmask) != k->cmask) continue;
                    apply_pair(k, re, im, i0);
                }
        }
    }
}

/* Uncontrolled cross-row gate: tiles of 256 columns, each tile walks all
   row pairs (transform_column_simd_tiled). */
static void run_cross_row(double* re, double* im, const gate_k* k, int col_bits, uint64_t rows, uint64_t row_size) {
    uint64_t rstride = k->stride >> col_bits;
    uint64_t tiles = (row_size + L2_TILE_COLS - 1) / L2_TILE_COLS;
#pragma omp parallel for schedule(static)
    for (uint64_t t = 0; t < tiles; ++t) {
        uint64_t c0 = t * L2_TILE_COLS, c1 = c0 + L2_TILE_COLS;
        if (c1 > row_size) c1 = row_size;
        for (uint64_t kk = 0; kk < rows / (2 * rstride); ++kk)
            for (uint64_t r = kk * 2 * rstride; r < kk * 2 * rstride + rstride; ++r)
                for (uint64_t c = c0; c < c1; ++c) apply_pair(k, re, im, r * row_size + c);
    }
}

/* Controlled / multi-controlled cross-row gates: one flat parallel pass over
   all pairs with the mask test. */
static void run_flat(double* re, double* im, const gate_k* k, uint64_t size) {
    uint64_t half = size / 2;
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < half; ++p) {
        uint64_t lo = p & (k->stride - 1);
        uint64_t i0 = ((p - lo) << 1) | lo;
        if ((i0 & k->cmask) != k->cmask) continue;
        apply_pair(k, re, im, i0);
    }
}

static uint64_t bitrev_n(uint64_t i, int n) {
    uint64_t r = 0;
    for (int b = 0; b < n; ++b) r |= ((i >> b) & 1ull) << (n - 1 - b);
    return r;
}

static int run_bitrev(double* re, double* im, int n, uint64_t mask) {
    uint64_t size = 1ull << n;
    if (mask == 0 || mask == size - 1) {
        double* sr = (double*)malloc(sizeof(double) * size);
        double* si = (double*)malloc(sizeof(double) * size);
        if (!sr || !si) { free(sr); free(si); return BFLY_ENOMEM; }
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < size; ++i) { uint64_t r = bitrev_n(i, n); sr[i] = re[r]; si[i] = im[r]; }
        memcpy(re, sr, sizeof(double) * size);
        memcpy(im, si, sizeof(double) * size);
        free(sr); free(si);
        return BFLY_OK;
    }
    int t[64], k = 0;
    for (int b = 0; b < n; ++b) if ((mask >> b) & 1) t[k++] = b;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < size; ++i) {
        uint64_t is = 0, js = 0;
        for (int b = 0; b < k; ++b) if ((i >> t[b]) & 1) is |= 1ull << b;
        for (int b = 0; b < k; ++b) if ((is >> b) & 1) js |= 1ull << (k - 1 - b);
        if (is < js) {
            uint64_t j = i;
            for (int b = 0; b < k; ++b) { if ((js >> b) & 1) j |= 1ull << t[b]; else j &= ~(1ull << t[b]); }
            double tr = re[i], ti = im[i];
            re[i] = re[j]; im[i] = im[j]; re[j] = tr; im[j] = ti;
        }
    }
    return BFLY_OK;
}

static int is_row_local(const bfly_op* op, int col_bits) {
    if (op->kind != BFLY_OP_GATE) return 0;
    if (__builtin_popcountll(op->controls) > 1) return 0;
    uint64_t lim = 1ull << col_bits;
    return op->target < col_bits && op->controls < lim;
}

/* Returns 0 or a BFLY_E* code.  GATE, BITREV and SWAP ops only; threads <= 0
   keeps the OpenMP default; slack < 0 means the reference default (3). */
int grid_fused_execute_f64(int n, double* re, double* im, const bfly_op* ops, uint64_t nops,
                           int slack, int min_bits, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    if (slack < 0) slack = 3;
    if (min_bits < 0) min_bits = 3;
    int col_bits = n - slack;
    if (col_bits < min_bits) col_bits = min_bits;
    if (col_bits > n) col_bits = n;
    uint64_t rows = 1ull << (n - col_bits), row_size = 1ull << col_bits, size = 1ull << n;
    gate_k* buf = (gate_k*)malloc(sizeof(gate_k) * (nops ? nops : 1));
    if (!buf) return BFLY_ENOMEM;
    int rc = BFLY_OK;
    uint64_t i = 0;
    while (i < nops && rc == BFLY_OK) {
        const bfly_op* op = &ops[i];
        if (op->kind == BFLY_OP_GATE && is_row_local(op, col_bits)) {
            int ng = 0;
            while (i < nops && is_row_local(&ops[i], col_bits)) make_gate(&ops[i++], &buf[ng++]);
            run_row_group(re, im, buf, ng, rows, row_size);
        } else if (op->kind == BFLY_OP_GATE) {
            make_gate(op, &buf[0]);
            if (op->controls == 0 && op->target >= col_bits) run_cross_row(re, im, &buf[0], col_bits, rows, row_size);
            else run_flat(re, im, &buf[0], size);
            ++i;
        } else if (op->kind == BFLY_OP_BITREV) {
            rc = run_bitrev(re, im, n, op->aux);
            ++i;
        } else if (op->kind == BFLY_OP_SWAP) {
            bfly_op cx;
            memset(&cx, 0, sizeof cx);
            cx.kind = BFLY_OP_GATE; cx.gate = BFLY_GATE_X;
            int a = op->target, b = op->arg_i;
            if (a != b) {
                int seq[3][2] = {{a, b}, {b, a}, {a, b}};
                for (int s = 0; s < 3; ++s) {
                    cx.controls = 1ull << seq[s][0]; cx.target = seq[s][1];
                    make_gate(&cx, &buf[0]);
                    run_flat(re, im, &buf[0], size);
                }
            }
            ++i;
        } else {
            rc = BFLY_EUNSUPPORTED;
        }
    }
    free(buf);
    return rc;
}

int grid_fused_max_threads(void) { return omp_get_max_threads(); }
