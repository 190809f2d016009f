// symhip_run_noalloc: symhip_run without any device allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The only device memory a program's launches use beyond its tables, inputs and outputs is
// its work words: the work-queue counters of persistent schedules (and, when profiling, the
// per-tile stamps).  symhip_run passes the handle's own buffer; this entry point passes a
// file-scope array instead, one per device, held until the process ends.  Host code never
// uses the array's name as an address: the device address comes from hipGetSymbolAddress,
// resolved once per device on the first call with that device current.
//
// Calls on one device share the array, which keeps the last call's contents: every call
// clears the words its program uses (with a fill kernel, not a memset, so the call can be
// captured into a graph) before its launches.  Calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "symhip_internal.hpp"

static __device__ __align__(256) uint32_t scratch[SYMHIP_SCRATCH_WORDS];

namespace {

__global__ void __launch_bounds__(256) symhipScratchClear(uint32_t* words, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) words[i] = 0u;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
uint32_t* g_scratchPtr[kMaxDevices] = {};

// Device address of the current device's scratch array.
int scratchAddress(uint32_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return symhipFail((int)e, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return symhipFail(SYMHIP_E_INVALID, "symhip_run_noalloc: device ordinal " + std::to_string(dev) + " beyond 63");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return symhipFail(e != hipSuccess ? (int)e : SYMHIP_E_INVALID,
                              std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchPtr[dev] = static_cast<uint32_t*>(ptr);
    }
    *out = g_scratchPtr[dev];
    return SYMHIP_SUCCESS;
}

}  // namespace

extern "C" int symhip_run_noalloc(symhip_t h, void* stream) {
    if (!h) return symhipFail(SYMHIP_E_INVALID, "symhip_run_noalloc: null handle");
    if (h->functions.size() != h->launches.size()) return symhipFail(SYMHIP_E_NOT_COMPILED, "symhip: program not compiled");
    if (h->numWork > SYMHIP_SCRATCH_WORDS)
        return symhipFail(SYMHIP_E_INVALID, "symhip_run_noalloc: the program uses " + std::to_string(h->numWork) +
                                                " work words, more than the " + std::to_string(SYMHIP_SCRATCH_WORDS) +
                                                " of the static scratch array (profiled launches add 8 per tile)");
    uint32_t* ptr = nullptr;
    int rc = scratchAddress(&ptr);
    if (rc) return rc;
    if (h->numWork) {
        const uint32_t n = (uint32_t)h->numWork;
        hipLaunchKernelGGL(symhipScratchClear, dim3((n + 255u) / 256u), dim3(256), 0, (hipStream_t)stream, ptr, n);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return symhipFail((int)e, std::string("symhip_run_noalloc: clearing the scratch words: ") +
                                                           hipGetErrorString(e));
    }
    return symhipLaunchAll(h, stream, ptr);
}
