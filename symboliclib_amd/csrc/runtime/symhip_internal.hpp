// Internal to the runtime (not part of the C ABI): the program handle and the launch helpers
// shared by symhip.cpp and the device code compiled ahead of time (the STREAM copy, the result
// gather of reference-form programs and the no-allocation entry point symhip_run_noalloc).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
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
    // reference argument form per launch: (x, c + off[0], x + off[1], p + off[2]) instead of
    // this runtime's eight arguments
    std::vector<char> refForm;
    std::vector<std::array<uint64_t, 3>> refOffsets;
    uint64_t numConst = 0, numIndex = 0, numOutpos = 0;
    bool gather = false;  // after the launches: result[j] = x[o[j]] for every result j
};

// Enqueue launch `idx` of `h` on `stream` with `work` as the kernels' work-word buffer.
int symhipLaunch(symhip_t h, uint32_t idx, void* stream, uint32_t* work);
// Enqueue every launch of `h`, then the result gather when it is on.
int symhipLaunchAll(symhip_t h, void* stream, uint32_t* work);
// Enqueue result[j] = x[o[j]], j < numResults (defined with the device code).
int symhipGather(symhip_t h, void* stream);
// Record `msg` as the calling thread's last error; returns `code`.
int symhipFail(int code, const std::string& msg);
