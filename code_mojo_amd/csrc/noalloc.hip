// noalloc.hip -- the allocation-free entry point of the C ABI.
//
// bfly_prob_mask_noalloc is bfly_prob_mask for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc): the caller passes the device arrays and a device
// address for the result, the call only enqueues the two reduction kernels
// (FP64 TwoSum per thread, fixed tree, fixed block count per size: the same
// kernels and the same bits as bfly_prob_mask), and the per-block partials
// live in `scratch`, a file-scope device array of fixed size.  Each GPU holds
// its own copy until the process ends; calls that overlap on one GPU share it
// and race.  No hipMalloc, hipFree, memset or host synchronisation.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>

#include "internal.h"

namespace {

// 2 doubles (TwoSum value and carry) per reduction block, at most 1024 blocks.
constexpr size_t kScratchDoubles = 2048;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_setup_mu;
std::atomic<double*> g_scratch[kMaxDevices] = {};

// The address of this device's copy of `scratch`, looked up once per device
// with the device current.  The array's name is a host-side stand-in: only
// the address from hipGetSymbolAddress may be handed to a kernel.
int device_scratch(int device, double** out) {
    if (device < 0 || device >= kMaxDevices)
        return bfly::fail(BFLY_EINVAL, "device " + std::to_string(device) + " beyond the supported range");
    double* p = g_scratch[device].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> lk(g_setup_mu);
        p = g_scratch[device].load(std::memory_order_relaxed);
        if (!p) {
            void* addr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return bfly::hip_fail(e, "hipGetSymbolAddress(scratch)");
            p = static_cast<double*>(addr);
            g_scratch[device].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return BFLY_OK;
}

}  // namespace

extern "C" int bfly_prob_mask_noalloc(const void* re, const void* im, int32_t num_qubits, int32_t dtype,
                                      uint64_t mask, uint64_t value, double* out, void* stream) {
    if (!re || !im || !out) return bfly::fail(BFLY_EINVAL, "null device pointer");
    if (num_qubits < 1 || num_qubits > bfly::kMaxQubits)
        return bfly::fail(BFLY_EINVAL, "qubit count out of range: " + std::to_string(num_qubits) + " (1.." +
                                           std::to_string(bfly::kMaxQubits) + ")");
    if (dtype != BFLY_C128 && dtype != BFLY_C64) return bfly::fail(BFLY_EINVAL, "unknown dtype");
    if (value & ~mask) return bfly::fail(BFLY_EINVAL, "value has bits outside mask");
    const uint64_t size = 1ull << num_qubits;
    const size_t need = 2ull * bfly::reduce_blocks(size);
    if (need > kScratchDoubles)
        return bfly::fail(BFLY_EINVAL, "reduction of 2^" + std::to_string(num_qubits) + " amplitudes needs " +
                                           std::to_string(need) + " scratch doubles, the static buffer holds " +
                                           std::to_string(kScratchDoubles));
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    double* part = nullptr;
    if (const int rc = device_scratch(device, &part)) return rc;
    HIP_TRY(bfly::launch_reduce_into(dtype, bfly::R_PROB_MASK, re, im, nullptr, nullptr, size, mask, value, part, out,
                                     static_cast<hipStream_t>(stream)),
            "reduction launch");
    return BFLY_OK;
}
