// jit.h -- tile passes compiled into straight-line kernels at plan time
// (BFLY_FLAG_JIT); see jit.cpp.
#pragma once
#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "bfly_plan.h"

namespace bfly {

struct JitCode;   // a compiled code object (process-wide cache entry)

// Generated source of one tile pass and its parameter array (in the
// state's precision: cd for complex128, cf for complex64).
struct JitSource {
    std::string text;
    std::vector<double> cd;
    std::vector<float> cf;
    int threads = 0;
    unsigned long long ntiles = 0;
    bool persistent = false;       // blocks walk the tiles (else one block per tile)
    bool queue = false;            // persistent blocks take tiles from a device counter
};

struct JitPass {
    std::shared_ptr<JitCode> code;
    size_t const_off = 0;          // byte offset of the pass's parameters in JitPlan::consts
    int threads = 0;
    unsigned long long ntiles = 0;
    bool persistent = false;
    bool queue = false;            // takes its tile counter as a fourth argument
    int qidx = -1;                 // its counter among the state's tile counters
    unsigned grid = 0;             // blocks per launch on the loaded device
    hipFunction_t fn = nullptr;
};

struct JitPlan {
    bool built = false;
    int compiled = 0;              // sources compiled (not found in the cache) when the plan was built
    int device = -1;               // device the kernels are loaded on
    std::vector<int> step_pass;    // per plan step: index into passes, or -1 (interpreted)
    std::vector<JitPass> passes;
    std::vector<char> consts;      // every pass's parameters, 16-byte aligned per pass
    void* d_consts = nullptr;      // device copy (owned by the plan)
    int nqueue = 0;                // passes with a tile counter (the executing state holds them)
};

bool jit_supported(const HostPlan& hp, const Step& st);
JitSource jit_generate(const HostPlan& hp, const Step& st);
// Generates and compiles (hipRTC, host only) every supported tile pass.
int jit_build(JitPlan& jp, const HostPlan& hp, std::string& err);
// Loads the kernels on `device` (the caller made it current) and sizes their grids.
int jit_load(JitPlan& jp, int device, std::string& err);
hipError_t jit_launch(const JitPass& ps, void* re, void* im, const void* consts, unsigned long long* counter,
                      hipStream_t stream);
int jit_cache_size();

}  // namespace bfly
