/*
 * bfly_hip.h -- C ABI of the MI355X statevector engine (libbfly_hip.so).
 *
 * This library is the GPU execution strategy of the Butterfly simulator
 * (learnqc/code_mojo).  A Butterfly host (the Mojo `ExecutorFn` shim, the
 * Python mirror in code_mojo_amd/, or any C caller) lowers a circuit's
 * transformation list into an array of `bfly_op` once per circuit and crosses
 * this boundary once, instead of dispatching one CPU kernel per
 * transformation.
 *
 * Conventions (all identical to the reference):
 *   - amplitudes are SoA: re[2^n], im[2^n]; qubit q is bit q of the basis
 *     index (LSB = qubit 0)                     butterfly/core/state.mojo:19-33
 *   - a 1-qubit gate maps (a0, a1) -> G (a0, a1) with a0 = amp[i] (target bit
 *     0) and a1 = amp[i + 2^t]     butterfly/core/transformations_scalar.mojo:10-40
 *   - a gate with controls acts where (index & controls) == controls
 *                                  butterfly/core/transformations_scalar.mojo:590
 *   - an m-qubit unitary acts on qubits [target, target+m), local index k
 *     <-> offset k << target, row-major, out = U in
 *                                  butterfly/core/transformations_unitary.mojo:22-75
 *   - BITREV reverses the order of the target bits (all bits when the mask is
 *     0): out[i] = in[rev(i)]           butterfly/core/state.mojo:131-204,252-301
 *
 * Every entry point returns 0 on success or a negative BFLY_E* code; the
 * message of the last failure on the calling thread is bfly_last_error().
 * A state handle owns its HBM copy and one HIP stream; calls on one handle
 * are ordered on that stream, calls on different handles may overlap.
 * Pointers passed in are borrowed for the duration of the call.
 */
#ifndef BFLY_HIP_H
#define BFLY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define BFLY_OK            0
#define BFLY_EINVAL       -1  /* bad argument (qubit out of range, size mismatch, ...) */
#define BFLY_ENOMEM       -2  /* device or host allocation failed / over budget */
#define BFLY_EHIP         -3  /* HIP runtime error */
#define BFLY_EUNSUPPORTED -4  /* transformation the GPU strategy cannot lower */
#define BFLY_ENUMERIC     -5  /* norm check failed (non-finite or drifted state) */
#define BFLY_ETIMEDOUT    -6  /* bfly_engine_wait: the request did not finish in time */
#define BFLY_ESTATE       -7  /* request in the wrong state for the call (not finished, cancelled, failed) */

/* ---- amplitude precision ---------------------------------------------- */
#define BFLY_C128 0           /* f64 re/im: reference default FloatType   types.mojo:7 */
#define BFLY_C64  1           /* f32 re/im: reference `-D FLOAT32` build  types.mojo:7 */

/* ---- op kinds ---------------------------------------------------------- */
#define BFLY_OP_GATE     0    /* 2x2 gate on `target`, controls mask `controls`  circuit.mojo:46-67 */
#define BFLY_OP_BITREV   1    /* reverse bits of mask `aux` (0 = all qubits)     quantum_circuit.mojo:99-136 */
#define BFLY_OP_SWAP     2    /* exchange qubits `target` and `arg_i`            state.mojo:304-317, executors.mojo:231-265 */
#define BFLY_OP_UNITARY  3    /* dense 2^m block, m = `arg_i`, on [target, target+m); matrix at udata[2*aux ..]
                                 (re,im interleaved, row-major); optional controls mask  transformations_unitary.mojo:22-139 */
#define BFLY_OP_PHASE_FLIP 4  /* negate every amplitude i with (i & aux) == controls (controls within aux):
                                 classical_oracle (aux = all, controls = the marked index)  algos/grover.mojo:64-66,
                                 apply_grover_inversion_0 (controls = 0)  algos/grover_classical.mojo:262-281 */
#define BFLY_OP_DIFFUSE  5    /* reflection about the uniform superposition of the bits in `aux`, per block of
                                 the other bits, with mean = (sum of the block) / 2^popcount(aux):
                                 arg_i = 0: a <- a - 2 mean, the action of H X MCP(pi) X H on those bits
                                            (diffuser_circuit, algos/grover.mojo:45-62);
                                 arg_i = 1: re <- 2 mean_re - re, im <- -2 mean_im - im, exactly
                                            apply_grover_diffuser (algos/grover_classical.mojo:284-326);
                                 at most 12 bits unless aux covers every qubit */
#define BFLY_OP_MEASURE  6    /* apply_measure (state.mojo:447-553): arg_i = number of targets, their list at
                                 udata[2*aux ..] as (qubit, value) pairs, value -1 = free, 0 / 1 = forced;
                                 gate = 1: seed in `controls` (int64 >= 0), gate = 0: seed from the clock;
                                 the plan synchronises on the state's stream at this op (bfly_measure) */

/* ---- gate kinds: the reference's GateKind values (gates.mojo:83-91) ---- */
#define BFLY_GATE_H      0
#define BFLY_GATE_P      1
#define BFLY_GATE_X      2
#define BFLY_GATE_Y      3
#define BFLY_GATE_Z      4
#define BFLY_GATE_RX     5
#define BFLY_GATE_RY     6
#define BFLY_GATE_RZ     7
#define BFLY_GATE_CUSTOM 8

/* One lowered transformation.  104 bytes, 8-byte aligned. */
typedef struct bfly_op {
    int32_t  kind;      /* BFLY_OP_* */
    int32_t  gate;      /* BFLY_GATE_* (GATE ops) */
    int32_t  target;    /* target qubit / first qubit of a block / qubit a of a swap */
    int32_t  arg_i;     /* SWAP: qubit b;  UNITARY: m */
    uint64_t controls;  /* control-qubit mask (GATE, UNITARY) */
    uint64_t aux;       /* BITREV: target mask;  UNITARY: complex offset into udata */
    double   theta;     /* angle for P / RX / RY / RZ (GateInfo.arg) */
    double   mat[8];    /* 2x2 gate, row-major, (re, im) per entry: g00 g01 g10 g11 */
} bfly_op;

/* ---- execution modes --------------------------------------------------- */
#define BFLY_MODE_FUSED   0   /* LDS tile passes: many gates per HBM sweep (default) */
#define BFLY_MODE_UNFUSED 1   /* one strided butterfly launch per gate */
#define BFLY_MODE_BLOCK   2   /* gates fused into dense 2^5 blocks on MFMA (FP32 for c64, FP64 for c128) */
#define BFLY_MODE_AUTO    3   /* build the FUSED and the BLOCK schedule, keep the one a cost model
                                 (measured pass times on MI355X) predicts faster */

typedef struct bfly_exec_opts {
    int32_t mode;            /* BFLY_MODE_* */
    int32_t tile_bits;       /* k: amplitudes per LDS tile = 2^k (0 = tuned default) */
    int32_t low_bits;        /* L: low qubits kept in every tile for coalescing (0 = default) */
    int32_t block_qubits;    /* BLOCK mode block width (0 = 5) */
    int32_t max_gates_per_pass; /* 0 = unlimited */
    int32_t renorm_every;    /* BLOCK mode: FP64 compensated norm check every N block passes (0 = off) */
    int32_t threads;         /* workgroup size for tile passes (0 = tuned default) */
    int32_t flags;           /* BFLY_FLAG_* */
} bfly_exec_opts;

#define BFLY_FLAG_SYNC      1 /* block until the device work is done before returning */
#define BFLY_FLAG_PROFILE   2 /* bracket every launch with HIP events (see bfly_profile_*) */
#define BFLY_FLAG_GENERIC_PAIRS 4 /* fused mode: apply every gate as its 2x2 product (no factoring of
                                     uncontrolled one-qubit gates into phases and a scaled rotation,
                                     no Grover peephole): the reference's per-gate arithmetic */
#define BFLY_FLAG_JIT       8 /* fused mode: compile every tile pass into its own straight-line kernel
                                 (hipRTC, at bfly_plan_create; cached per process by pass structure);
                                 same amplitudes as the interpreting tile kernel.  Worth it for plans
                                 executed many times: compiling takes about 0.3 s per new pass */
#define BFLY_FLAG_FIRST_FIT 16 /* fused mode: each tile pass takes its qubits in program order (first
                                  fit) instead of searching for the set that carries the most gates
                                  (faster planning, usually more passes) */

typedef struct bfly_state bfly_state;   /* opaque: device-resident SoA state */
typedef struct bfly_plan  bfly_plan;    /* opaque: compiled pass schedule    */

/* ---- library ----------------------------------------------------------- */
/* Thread-local message describing the last failure (empty when none). */
const char* bfly_last_error(void);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int bfly_version(void);
/* Number of visible devices (0 when no GPU); never fails. */
int bfly_device_count(void);

/* ---- state: replaces State / QuantumState (state.mojo:19-98) ----------- */
/* Allocate a 2^n-amplitude state on the current device, initialised to |0..0>
   (state.mojo:26-33).  Budget errors return BFLY_ENOMEM. */
int bfly_state_create(bfly_state** out, int32_t num_qubits, int32_t dtype);
int bfly_state_destroy(bfly_state* s);
/* Re-initialise to |0..0>. */
int bfly_state_reset(bfly_state* s);
/* Copy host SoA arrays (len = 2^n elements of the state's dtype) in / out. */
int bfly_state_upload(bfly_state* s, const void* re, const void* im, uint64_t len);
int bfly_state_download(bfly_state* s, void* re, void* im, uint64_t len);
/* Copy one state into another of the same qubit count on the same device (device to
   device; converts complex64 <-> complex128).  Asynchronous on dst's stream, ordered after
   src's pending work; src's later work waits for the copy, so the copy always reads src as
   it was at the call.  dst == src is rejected. */
int bfly_state_copy(bfly_state* dst, const bfly_state* src);
int bfly_state_info(const bfly_state* s, int32_t* num_qubits, int32_t* dtype);
/* Device pointers of the re / im arrays (for callers that share memory). */
int bfly_state_device_ptrs(bfly_state* s, void** re, void** im);
/* Wait for all work queued on the state's stream. */
int bfly_state_synchronize(bfly_state* s);

/* ---- execution: replaces execute() / ExecutorFn (executors.mojo:611-613,723-730) */
/* Compile a pass schedule for an n-qubit circuit (host only, no device work). */
int bfly_plan_create(bfly_plan** out, int32_t num_qubits, int32_t dtype,
                     const bfly_op* ops, uint64_t nops,
                     const double* udata, uint64_t udata_len,
                     const bfly_exec_opts* opts);
int bfly_plan_destroy(bfly_plan* p);
/* Counts: passes = HBM sweeps over the state (tile passes + permutations +
   dense blocks), launches = kernel launches, gates = lowered gate ops. */
int bfly_plan_stats(const bfly_plan* p, int64_t* passes, int64_t* launches, int64_t* gates);
/* Text description of the compiled schedule (one record per line; floats as
   C99 hex literals, so the schedule can be replayed exactly).  Writes at most
   `len` bytes including the terminating NUL and stores the full size
   (including the NUL) in *needed.  No counterpart in the reference: the
   schedule replaces its per-transformation dispatch. */
int bfly_plan_dump(const bfly_plan* p, char* buf, uint64_t len, uint64_t* needed);
/* BFLY_FLAG_JIT plans: tile passes that run as their own generated kernel, and how many of
   those were compiled when the plan was created (the others came from the process-wide
   cache).  Plans built without the flag report 0 / 0.  No counterpart in the reference
   (plan introspection, like bfly_plan_stats). */
int bfly_plan_jit_info(const bfly_plan* p, int64_t* kernels, int64_t* compiled);
/* Run a compiled plan on a state (asynchronous unless BFLY_FLAG_SYNC was set). */
int bfly_plan_execute(bfly_plan* p, bfly_state* s);
/* ---- autotuner: tile shape and fusion width ------------------------------
   No counterpart in the reference (its strategies take fixed layouts:
   compute_grid_layout, executors.mojo:213-228).  Candidate schedules are
   the base options plus, for each mode whose bit (1 << BFLY_MODE_*) is set in
   `modes`: FUSED at the default tile size and one size either side (and, for
   a state no wider than the default tile, one and two bits narrower than the
   state), each with no cap and with a cap of 2k gates per pass (k = tile
   bits); BLOCK; AUTO.  Options that compile to an identical schedule are kept
   once. */
#define BFLY_TUNE_ALL ((1 << BFLY_MODE_FUSED) | (1 << BFLY_MODE_BLOCK) | (1 << BFLY_MODE_AUTO))
/* Host only: the candidate list for an n-qubit circuit (at most `cap`
   written; *count = how many there are). */
int bfly_autotune_candidates(int32_t num_qubits, int32_t dtype, const bfly_exec_opts* base, int32_t modes,
                             bfly_exec_opts* out, int32_t cap, int32_t* count);
/* Times every candidate's schedule of this op list on a scratch state of the
   circuit's size (allocated and freed here; |0..0> start, one untimed run,
   then the median of `repeats` runs bracketed by HIP events on the state's
   stream) and returns the fastest options in *best, its time in *best_ms and
   the base options' time in *base_ms.  Candidates that fail to compile are
   skipped. */
int bfly_autotune(int32_t num_qubits, int32_t dtype, const bfly_op* ops, uint64_t nops,
                  const double* udata, uint64_t udata_len, const bfly_exec_opts* base, int32_t modes,
                  int32_t repeats, bfly_exec_opts* best, double* best_ms, double* base_ms);
/* Plan + execute + destroy in one call. */
int bfly_apply(bfly_state* s, const bfly_op* ops, uint64_t nops,
               const double* udata, uint64_t udata_len, const bfly_exec_opts* opts);
/* One-shot host strategy for an ExecutorFn: upload re/im, run ops, download
   back into re/im in place (len = 2^n). */
int bfly_execute_host(int32_t num_qubits, int32_t dtype, void* re, void* im, uint64_t len,
                      const bfly_op* ops, uint64_t nops,
                      const double* udata, uint64_t udata_len, const bfly_exec_opts* opts);

/* ---- batched execution: many circuits, one launch per round ------------
   No counterpart in the reference (it runs one circuit at a time).  Round r
   runs step r of every plan; the tile passes of one round that share a tile
   size are one persistent launch.  States and plans are borrowed and must
   outlive the batch; each state gets its own plan. */
typedef struct bfly_batch bfly_batch;
#define BFLY_BATCH_GRAPH 1    /* capture every round into one hipGraph, replayed by bfly_batch_launch
                                 (not under a rocprofiler-sdk tool, e.g. rocprofv3: there the same
                                 kernels launch directly on the batch's stream); at most 2048
                                 launches, BFLY_EUNSUPPORTED beyond */
int bfly_batch_create(bfly_batch** out, bfly_state* const* states, bfly_plan* const* plans, int32_t count,
                      int32_t flags);
/* Enqueue all rounds (ordered after pending work on every state's stream;
   later calls on those states are ordered after the batch). */
int bfly_batch_launch(bfly_batch* b);
int bfly_batch_synchronize(bfly_batch* b);
int bfly_batch_stats(const bfly_batch* b, int64_t* rounds, int64_t* launches);
int bfly_batch_destroy(bfly_batch* b);

/* ---- serving engine: continuous batching with HBM-budget admission ------
   No counterpart in the reference.  Requests (one circuit + initial state
   each) are admitted first-fit against a budget of 16*2^n bytes (complex128)
   or 8*2^n bytes (complex64) per resident state; admitted requests keep their
   state in HBM; a worker thread issues rounds in which the next step of every
   small running request joins one batched launch per tile size, while large
   requests (n >= large_qubits) run their whole plan on their own stream and
   overlap with the rounds.  Every result equals the request's result alone. */
typedef struct bfly_engine bfly_engine;
typedef struct bfly_engine_opts {
    uint64_t budget_bytes;        /* 0 = 90 % of the device memory free at creation */
    int32_t max_inflight_rounds;  /* rounds enqueued ahead of completion (0 = 3) */
    int32_t large_qubits;         /* requests with n >= this run on their own stream, two at a time (0 = 21) */
    int32_t max_round_requests;   /* requests per round (0 = unlimited) */
    int32_t flags;
    int32_t planner_threads;      /* host threads compiling submitted circuits (0 = 8) */
    int32_t reserved;
} bfly_engine_opts;

#define BFLY_INIT_ZERO   0        /* |0..0> */
#define BFLY_INIT_HOST   1        /* init_re / init_im, copied at submit */
#define BFLY_INIT_RANDOM 2        /* re, im ~ N(0,1) from init_seed (SplitMix64 + Box-Muller), normalised */

typedef struct bfly_request {
    int32_t num_qubits;
    int32_t dtype;
    const bfly_op* ops;
    uint64_t nops;
    const double* udata;
    uint64_t udata_len;
    const void* init_re;
    const void* init_im;
    int32_t init_kind;            /* BFLY_INIT_* */
    int32_t pad0;
    uint64_t init_seed;
    bfly_exec_opts exec;          /* planning options of this request */
} bfly_request;

#define BFLY_REQ_QUEUED    0
#define BFLY_REQ_RUNNING   1
#define BFLY_REQ_DONE      2
#define BFLY_REQ_FAILED    3
#define BFLY_REQ_CANCELLED 4

int bfly_engine_create(bfly_engine** out, const bfly_engine_opts* opts);
/* Plans the circuit (caller's thread) and queues it; *id names it afterwards.
   A state larger than the whole budget is refused with BFLY_ENOMEM. */
int bfly_engine_submit(bfly_engine* e, const bfly_request* r, uint64_t* id);
/* Status and timestamps (seconds since engine creation; 0 when not reached). */
int bfly_engine_status(bfly_engine* e, uint64_t id, int32_t* status, double* t_submit, double* t_admit,
                       double* t_done);
/* Block until the request is DONE / FAILED / CANCELLED or timeout_s passes
   (BFLY_ETIMEDOUT); timeout_s < 0 waits forever. */
int bfly_engine_wait(bfly_engine* e, uint64_t id, double timeout_s);
/* Copy a DONE request's final state to the host and free its HBM. */
int bfly_engine_fetch(bfly_engine* e, uint64_t id, void* re, void* im, uint64_t len);
/* Copy amplitudes [first, first + count) of a DONE request's state to the host (re, im:
   count elements of the request's precision each); the request stays in the engine until
   fetched or released.  For results too large to copy whole (a 33-qubit complex128 state is
   128 GiB).  No counterpart in the reference (its states live in host memory). */
int bfly_engine_fetch_range(bfly_engine* e, uint64_t id, uint64_t first, uint64_t count, void* re, void* im);
/* Free a finished request without copying it out. */
int bfly_engine_release(bfly_engine* e, uint64_t id);
/* Queued: dropped at once; running: no further steps are issued. */
int bfly_engine_cancel(bfly_engine* e, uint64_t id);
int bfly_engine_stats(bfly_engine* e, int64_t* admitted, int64_t* completed, int64_t* rounds, int64_t* launches,
                      uint64_t* bytes_in_use, uint64_t* peak_bytes);
/* Message of a FAILED request. */
int bfly_engine_error(bfly_engine* e, uint64_t id, char* buf, uint64_t len);
int bfly_engine_destroy(bfly_engine* e);

/* Fill a state with re, im ~ N(0,1) from `seed` (SplitMix64 counter stream,
   Box-Muller in FP64) and normalise it with the compensated norm. */
int bfly_state_fill_random(bfly_state* s, uint64_t seed);

/* ---- permutations: replaces bit_reverse_state / partial_bit_reverse_state
        (state.mojo:131-204,252-301) and swap_bits_state (state.mojo:304-317) -- */
int bfly_bitrev(bfly_state* s, uint64_t target_mask);
int bfly_swap_qubits(bfly_state* s, int32_t a, int32_t b);

/* ---- reductions: replaces probability_on_predicate / inner
        (grover_classical.mojo:8-31,109-131) and verify_states_equal
        (benchmark_verify.mojo:15-67); FP64 compensated, fixed order ---- */
int bfly_norm2(bfly_state* s, double* out);                       /* sum |a|^2 */
int bfly_prob_mask(bfly_state* s, uint64_t mask, uint64_t value, double* out); /* sum over (i & mask) == value */
/* Allocation-free variant of bfly_prob_mask.  No counterpart in the reference: it is for
   callers that cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  re, im: device arrays of 2^num_qubits amplitudes
   of `dtype`; out: a DEVICE address receiving one double, sum |a_i|^2 over (i & mask) ==
   value (the same kernels and bits as bfly_prob_mask); stream: a hipStream_t (NULL = the
   null stream).  The call enqueues two kernels and returns: no hipMalloc / hipFree, no
   memset, no host synchronisation.  The per-block partials live in a static device array of
   2048 doubles (16 KiB) that each GPU holds until the process ends; every call on a GPU
   shares it, so calls that overlap on one GPU race.  The first call on a device looks up
   that array's address (hipGetSymbolAddress; no allocation).  num_qubits 1..40. */
int bfly_prob_mask_noalloc(const void* re, const void* im, int32_t num_qubits, int32_t dtype, uint64_t mask,
                           uint64_t value, double* out, void* stream);
int bfly_marginal(bfly_state* s, const int32_t* targets, int32_t m, double* out); /* 2^m probabilities */
int bfly_inner(bfly_state* a, bfly_state* b, double* re, double* im);          /* <a|b> */
int bfly_diff_norm2(bfly_state* a, bfly_state* b, double* out);                /* sum |a-b|^2 */
/* Multiply every amplitude by `scale` (renormalisation). */
int bfly_scale(bfly_state* s, double scale);

/* ---- measurement and sampling: replaces apply_measure (state.mojo:447-553) --
   targets[k] with values[k] = -1 (free) or 0 / 1 (forced).  Outcome
   probabilities of the free targets (bit k of an outcome = the k-th free
   target) among the amplitudes matching the forced values are FP64
   compensated sums on the device; the outcome is chosen on the host as the
   reference does: one MINSTD step rng = (seed mod 2147483647) * 48271 mod
   2147483647, r = rng / 2147483647 * total in the state's precision, the
   first outcome whose running sum reaches r.  The state is collapsed onto it
   and divided by sqrt(its probability).  Up to 24 free targets the 2^f
   outcome sums are formed; with more (up to 64) the outcome is found level by
   level over the distinct free qubits (at most 2^24 bucket sums per level,
   one pass over the state each) with FP64 running sums and r for both
   precisions, since a float32 running sum over 2^25 or more outcomes
   stagnates.  seed < 0 takes the clock, as the reference does
   without a configured measure_seed.  `outcome` / `prob` (optional) receive
   the chosen outcome and its (unnormalised) probability.  Synchronous. */
int bfly_measure(bfly_state* s, const int32_t* targets, const int32_t* values, int32_t m, int64_t seed,
                 uint64_t* outcome, double* prob);
/* `shots` outcomes of `targets` (all free) without collapse: shot j uses the
   (j+1)-th state of the same MINSTD sequence (shot 0 is the outcome
   bfly_measure would choose with `seed`).  seed >= 0, not a multiple of
   2147483647, at most 24 targets.  Synchronous; the state is unchanged. */
int bfly_sample(bfly_state* s, const int32_t* targets, int32_t m, int64_t seed, uint64_t shots, uint64_t* samples);

/* ---- circuit ingestion: replaces parse_qasm3_string (utils/qasm3_import.mojo:52-123)
   and the lowering of its circuit.  The reference's statements (qubit[n], h, x,
   p(a), cp(a); angles pi, c*pi, pi/d, c/pi, numbers) keep its results and
   errors; also accepted: named registers, y z s sdg t tdg rx ry rz u3 U u2 u1
   phase cx cy cz crx cry crz ccx cphase cu1 swap mcp, ctrl(k) @, measure,
   barrier / bit declarations, several statements per line, arithmetic angle
   expressions.  Call with ops = NULL (and udata = NULL) to get the counts. */
int bfly_qasm3_lower(const char* text, int32_t* num_qubits, bfly_op* ops, uint64_t ops_cap, uint64_t* nops,
                     double* udata, uint64_t udata_cap, uint64_t* udata_len);

/* ---- profiling (BFLY_FLAG_PROFILE) -------------------------------------- */
/* Per kernel family: launches and summed device milliseconds since reset.
   `name` receives a static string.  Returns BFLY_EINVAL past the last family. */
int bfly_profile_reset(void);
int bfly_profile_get(int32_t index, const char** name, int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* BFLY_HIP_H */
