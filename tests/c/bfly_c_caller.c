/* bfly_c_caller.c -- a C host of the GPU strategy, compiled against
 * include/bfly_hip.h alone and linked to libbfly_hip.so, the way a Butterfly
 * host written in C (or bound over an FFI) calls it.
 *
 * It lowers a 12-qubit circuit by hand into bfly_op records (gates with and
 * without controls, a custom 2x2, a SWAP, a full BITREV, a controlled 2-qubit
 * unitary from udata, a phase flip and the classical Grover diffuser) and runs
 * it five ways:
 *   0  bfly_execute_host, complex128, default options
 *   1  bfly_state_create / bfly_plan_create (BFLY_FLAG_JIT) / bfly_plan_execute
 *      / bfly_state_download, complex128
 *   2  the same handles with BFLY_MODE_UNFUSED (one launch per gate)
 *   3  bfly_execute_host, complex64 (float arrays)
 *   4  bfly_state_upload of a host state, bfly_apply, complex128
 * and checks the error paths (bad target, bad qubit count, bad dtype) return
 * BFLY_EINVAL with a message.  argv[1] names a binary file that receives the
 * op array, udata, the start state of run 4 and the five results; the test
 * that runs this program compares them with the oracle.
 *
 * Exit status 0 = every call returned BFLY_OK where it should and the codes
 * of the error paths were right; the amplitudes are judged by the caller.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bfly_hip.h"

_Static_assert(sizeof(bfly_op) == 104, "bfly_op is 104 bytes");
_Static_assert(sizeof(bfly_exec_opts) == 32, "bfly_exec_opts is eight int32");

#define NQ 12
#define LEN (1u << NQ)
#define MAXOPS 256
#define NRUNS 5

static bfly_op ops[MAXOPS];
static uint64_t nops;
static double udata[64];
static uint64_t nud;        /* complex entries */

static int check(int rc, const char* what) {
    if (rc != BFLY_OK) {
        fprintf(stderr, "c_caller: %s failed (%d): %s\n", what, rc, bfly_last_error());
        exit(1);
    }
    return rc;
}

static bfly_op* push(int kind) {
    if (nops >= MAXOPS) {
        fprintf(stderr, "c_caller: op array full\n");
        exit(1);
    }
    bfly_op* o = &ops[nops++];
    memset(o, 0, sizeof *o);
    o->kind = kind;
    return o;
}

/* a GATE record: kind, target, control mask, angle, 2x2 (re, im) row-major */
static void gate(int g, int target, uint64_t controls, double theta, const double m[8]) {
    bfly_op* o = push(BFLY_OP_GATE);
    o->gate = g;
    o->target = target;
    o->controls = controls;
    o->theta = theta;
    memcpy(o->mat, m, sizeof o->mat);
}

static void build_circuit(void) {
    const double r = sqrt(0.5);
    const double H[8] = {r, 0, r, 0, r, 0, -r, 0};
    const double X[8] = {0, 0, 1, 0, 1, 0, 0, 0};
    for (int q = 0; q < NQ; ++q) gate(BFLY_GATE_H, q, 0, 0.0, H);
    for (int q = 0; q + 1 < NQ; q += 2) gate(BFLY_GATE_X, q + 1, 1ull << q, 0.0, X);
    for (int q = 0; q < NQ; ++q) {            /* RZ(t) = diag(e^{-it/2}, e^{it/2}) */
        const double t = 0.3 + 0.17 * q;
        const double rz[8] = {cos(t / 2), -sin(t / 2), 0, 0, 0, 0, cos(t / 2), sin(t / 2)};
        gate(BFLY_GATE_RZ, q, 0, t, rz);
    }
    {                                         /* RY(0.9) on every odd qubit */
        const double t = 0.9;
        const double ry[8] = {cos(t / 2), 0, -sin(t / 2), 0, sin(t / 2), 0, cos(t / 2), 0};
        for (int q = 1; q < NQ; q += 2) gate(BFLY_GATE_RY, q, 0, t, ry);
    }
    {                                         /* doubly controlled P(0.7) */
        const double t = 0.7;
        const double p[8] = {1, 0, 0, 0, 0, 0, cos(t), sin(t)};
        gate(BFLY_GATE_P, 5, (1ull << 2) | (1ull << 9), t, p);
    }
    {                                         /* custom U3(1.1, 0.4, -0.8) with a control */
        const double th = 1.1, ph = 0.4, la = -0.8;
        const double c = cos(th / 2), s = sin(th / 2);
        const double u[8] = {c, 0, -s * cos(la), -s * sin(la), s * cos(ph), s * sin(ph),
                             c * cos(ph + la), c * sin(ph + la)};
        gate(BFLY_GATE_CUSTOM, 7, 1ull << 0, 0.0, u);
        gate(BFLY_GATE_CUSTOM, 11, 0, 0.0, u);
    }
    bfly_op* sw = push(BFLY_OP_SWAP);         /* exchange qubits 2 and 9 */
    sw->target = 2;
    sw->arg_i = 9;
    push(BFLY_OP_BITREV);                     /* aux 0 = every qubit */
    {                                         /* 2-qubit DFT on [3, 5), controlled by qubit 10 */
        bfly_op* u = push(BFLY_OP_UNITARY);
        u->target = 3;
        u->arg_i = 2;
        u->controls = 1ull << 10;
        u->aux = nud;
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 4; ++k) {
                const double a = 2.0 * M_PI * (double)(i * k) / 4.0;
                udata[2 * nud] = 0.5 * cos(a);
                udata[2 * nud + 1] = 0.5 * sin(a);
                ++nud;
            }
    }
    bfly_op* pf = push(BFLY_OP_PHASE_FLIP);   /* negate amplitude 1234 */
    pf->aux = LEN - 1;
    pf->controls = 1234;
    bfly_op* df = push(BFLY_OP_DIFFUSE);      /* the reference's classical diffuser, all qubits */
    df->aux = LEN - 1;
    df->arg_i = 1;
    for (int q = 0; q < NQ; q += 3) gate(BFLY_GATE_H, q, 0, 0.0, H);
}

static double out_re[NRUNS][LEN], out_im[NRUNS][LEN];
static double start_re[LEN], start_im[LEN];

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s OUTPUT\n", argv[0]);
        return 2;
    }
    if (bfly_device_count() < 1) {
        fprintf(stderr, "c_caller: no GPU visible\n");
        return 1;
    }
    build_circuit();
    printf("c_caller: library version %d, %llu ops, %llu udata entries\n", bfly_version(),
           (unsigned long long)nops, (unsigned long long)nud);

    /* 0: one-shot host strategy, complex128 */
    out_re[0][0] = 1.0;
    check(bfly_execute_host(NQ, BFLY_C128, out_re[0], out_im[0], LEN, ops, nops, udata, nud, NULL),
          "bfly_execute_host c128");

    /* 1 and 2: resident state and compiled plans */
    for (int run = 1; run <= 2; ++run) {
        bfly_exec_opts o;
        memset(&o, 0, sizeof o);
        if (run == 1) o.flags = BFLY_FLAG_JIT;
        else o.mode = BFLY_MODE_UNFUSED;
        bfly_state* s = NULL;
        bfly_plan* p = NULL;
        check(bfly_state_create(&s, NQ, BFLY_C128), "bfly_state_create");
        check(bfly_plan_create(&p, NQ, BFLY_C128, ops, nops, udata, nud, &o), "bfly_plan_create");
        int64_t passes = 0, launches = 0, gates = 0;
        check(bfly_plan_stats(p, &passes, &launches, &gates), "bfly_plan_stats");
        check(bfly_plan_execute(p, s), "bfly_plan_execute");
        double norm = 0.0;
        check(bfly_norm2(s, &norm), "bfly_norm2");
        check(bfly_state_download(s, out_re[run], out_im[run], LEN), "bfly_state_download");
        printf("c_caller: run %d (%s): %lld passes, %lld launches, %lld gates, norm^2 %.17g\n", run,
               run == 1 ? "fused, generated kernels" : "unfused", (long long)passes, (long long)launches,
               (long long)gates, norm);
        /* the reference's classical diffuser (im <- -2 mean_im - im) does not
           keep the norm of a complex state: compare the device's compensated
           norm with the host's sum over the downloaded amplitudes instead */
        double host = 0.0;
        for (uint32_t i = 0; i < LEN; ++i) host += out_re[run][i] * out_re[run][i] + out_im[run][i] * out_im[run][i];
        if (fabs(norm - host) > 1e-12 * host) {
            fprintf(stderr, "c_caller: device norm^2 %.17g, host %.17g\n", norm, host);
            return 1;
        }
        check(bfly_plan_destroy(p), "bfly_plan_destroy");
        check(bfly_state_destroy(s), "bfly_state_destroy");
    }

    /* 3: complex64 through the one-shot strategy (float arrays) */
    {
        static float re32[LEN], im32[LEN];
        re32[0] = 1.0f;
        check(bfly_execute_host(NQ, BFLY_C64, re32, im32, LEN, ops, nops, udata, nud, NULL),
              "bfly_execute_host c64");
        for (uint32_t i = 0; i < LEN; ++i) {
            out_re[3][i] = re32[i];
            out_im[3][i] = im32[i];
        }
    }

    /* 4: a host start state uploaded into a resident state, bfly_apply */
    {
        double nrm = 0.0;
        for (uint32_t i = 0; i < LEN; ++i) {
            start_re[i] = cos(0.37 * i) / (1.0 + (i % 7));
            start_im[i] = sin(0.11 * i + 0.5) / (1.0 + (i % 5));
            nrm += start_re[i] * start_re[i] + start_im[i] * start_im[i];
        }
        nrm = 1.0 / sqrt(nrm);
        for (uint32_t i = 0; i < LEN; ++i) {
            start_re[i] *= nrm;
            start_im[i] *= nrm;
        }
        bfly_state* s = NULL;
        check(bfly_state_create(&s, NQ, BFLY_C128), "bfly_state_create");
        check(bfly_state_upload(s, start_re, start_im, LEN), "bfly_state_upload");
        check(bfly_apply(s, ops, nops, udata, nud, NULL), "bfly_apply");
        check(bfly_state_synchronize(s), "bfly_state_synchronize");
        check(bfly_state_download(s, out_re[4], out_im[4], LEN), "bfly_state_download");
        check(bfly_state_destroy(s), "bfly_state_destroy");
    }

    /* error paths: codes and messages, no device work */
    {
        bfly_plan* p = NULL;
        bfly_op bad = ops[0];
        bad.target = NQ;                      /* out of range */
        int rc = bfly_plan_create(&p, NQ, BFLY_C128, &bad, 1, NULL, 0, NULL);
        if (rc != BFLY_EINVAL || strlen(bfly_last_error()) == 0) {
            fprintf(stderr, "c_caller: bad target gave %d '%s'\n", rc, bfly_last_error());
            return 1;
        }
        printf("c_caller: bad target -> %d: %s\n", rc, bfly_last_error());
        bfly_state* s = NULL;
        rc = bfly_state_create(&s, 0, BFLY_C128);
        if (rc != BFLY_EINVAL) {
            fprintf(stderr, "c_caller: zero qubits gave %d\n", rc);
            return 1;
        }
        rc = bfly_state_create(&s, NQ, 7);
        if (rc != BFLY_EINVAL) {
            fprintf(stderr, "c_caller: dtype 7 gave %d\n", rc);
            return 1;
        }
        printf("c_caller: error paths ok\n");
    }

    FILE* f = fopen(argv[1], "wb");
    if (!f) {
        perror("c_caller: open output");
        return 1;
    }
    const int32_t hdr[4] = {NQ, (int32_t)nops, (int32_t)nud, NRUNS};
    int ok = fwrite(hdr, sizeof hdr, 1, f) == 1 && fwrite(ops, sizeof(bfly_op), nops, f) == nops &&
             fwrite(udata, 2 * sizeof(double), nud, f) == nud && fwrite(start_re, sizeof start_re, 1, f) == 1 &&
             fwrite(start_im, sizeof start_im, 1, f) == 1;
    for (int r = 0; r < NRUNS && ok; ++r)
        ok = fwrite(out_re[r], sizeof out_re[r], 1, f) == 1 && fwrite(out_im[r], sizeof out_im[r], 1, f) == 1;
    if (fclose(f) != 0 || !ok) {
        fprintf(stderr, "c_caller: writing %s failed\n", argv[1]);
        return 1;
    }
    printf("c_caller: done\n");
    return 0;
}
