// internal.h -- objects behind the C ABI's opaque handles, shared by the
// translation units of libbfly_hip.so (api.hip, engine.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "bfly_plan.h"
#include "jit.h"
#include "kernels.h"

#include <vector>

struct bfly_state {
    int32_t n = 0;
    int32_t dtype = BFLY_C128;
    int device = 0;
    void* re = nullptr;
    void* im = nullptr;
    double* scratch = nullptr;   // reduction partials
    hipStream_t stream = nullptr;
    // tile counters of work-queue tile passes (one per such pass of the plan
    // being executed; calls on one state are ordered on its stream)
    unsigned long long* tiles = nullptr;
    int tiles_cap = 0;
};

struct bfly_plan {
    bfly::HostPlan hp;
    int device = -1;
    bfly::TileSweep* d_sweeps = nullptr;
    bfly::TileInstr* d_instrs = nullptr;
    double* d_udata = nullptr;
    float* d_bmat = nullptr;             // BLOCK mode block matrices
    double* d_btab = nullptr;              // BLOCK mode column factors and outer angles
    bool borrowed = false;   // device records live in a buffer the plan does not own
    bfly::JitPlan* jit = nullptr;   // BFLY_FLAG_JIT: generated tile-pass kernels
};

namespace bfly {

// Record `msg` as the calling thread's last error and return `code`.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int check_state(const bfly_state* s);
// Copy a plan's sweep / instruction / unitary records to `device` (once).
int upload_plan(bfly_plan* p, int device);
// Device bytes of a plan's records, and their asynchronous upload into a
// caller-owned buffer of at least that size (the host records must stay
// alive until the copy has run on `stream`).
size_t plan_device_bytes(const bfly_plan* p);
// Seeded N(0,1) state normalised to 1, enqueued on `stream`.
int fill_random_on(bfly_state* s, uint64_t seed, hipStream_t stream);
// apply_measure on a device state (synchronous; m = 0 does nothing).
int measure_state(bfly_state* s, const int32_t* targets, const int32_t* values, int32_t m, int64_t seed,
                  uint64_t* outcome, double* prob);
int upload_plan_into(bfly_plan* p, int device, void* buf, hipStream_t stream);
// Enqueue one plan step on `stream`.
hipError_t launch_step(const Step& st, bfly_state* s, bfly_plan* p, hipStream_t stream);

}  // namespace bfly

#define HIP_TRY(expr, what)                                      \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return bfly::hip_fail(_e, what);   \
    } while (0)
