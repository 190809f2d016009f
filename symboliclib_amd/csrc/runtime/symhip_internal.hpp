// Internal to the runtime (not part of the C ABI): the program handle and the launch helper
// shared by symhip.cpp and the device code compiled ahead of time (the STREAM copy and the
// no-allocation entry point symhip_run_noalloc).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "symhip.h"

struct symhip_program {
    std::vector<uint64_t> argStart, argLength, outStart, outLength;
    uint64_t numValues = 0, numResults = 0;
    double* x = nullptr;
    double* result = nullptr;
    double* c = nullptr;
    uint32_t* p = nullptr;
    uint32_t* o = nullptr;
    uint32_t* tiles = nullptr;
    uint32_t* work = nullptr;
    uint64_t numTilesData = 0, numWork = 0;
    std::vector<symhip_launch> launches;
    std::vector<uint32_t> grids;  // resolved workgroup counts
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::vector<std::string> names;
    hipModule_t module = nullptr;
    std::vector<hipFunction_t> functions;
};

// Enqueue launch `idx` of `h` on `stream` with `work` as the kernels' work-word buffer.
int symhipLaunch(symhip_t h, uint32_t idx, void* stream, uint32_t* work);
// Record `msg` as the calling thread's last error; returns `code`.
int symhipFail(int code, const std::string& msg);
